// Host side of the gfx950 G1 MSM backend: per-device context, launch
// sequence of the device stages, the latency-bound window fold, and the C ABI
// declared in include/blst_hip.h.
//
// The drop-in entry points keep the reference's checks and their order
// (blst/c_api.cpp:1382-1418, 1546-1611): they are the only part of the
// reference's control flow that callers can observe.  Everything past the
// checks is this backend's own pipeline.
#include "msm_kernels.hip.h"
#include "host_ec.h"
#include "../../include/blst_hip.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace bls;

namespace {

thread_local std::string g_err;
thread_local double g_phase[7];

struct HipFail {
    std::string msg;
};

#define HIPCK(expr)                                                                   \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw HipFail{std::string(#expr) + ": " + hipGetErrorString(e_)};         \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <class T> T* as() const { return static_cast<T*>(p); }
    void ensure(size_t bytes) {
        if (bytes <= cap && p) return;
        if (p) HIPCK(hipFree(p));
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 8, 4096);
        HIPCK(hipMalloc(&p, want));
        cap = want;
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    DevBuf pts, scl, ent, rank, hist, off, skey, sval, buckets, slot_key, slots, H, scan_tmp, dbg;
    DevBuf gtable, gen_tmp, gen_out;     // comb table of G, fixed-base scratch
    bool gtable_ready = false;
    hipEvent_t ev[8] = {};
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
    HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIPCK(hipEventCreate(&e));
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

struct Plan {
    uint32_t n, nbits, nbytes, c, nwin, B, bit_base;
    bool fold;        // combine windows (full MSM) or return the single window (tile)
};

// ---------------------------------------------------------------- device

void scan_exclusive(Ctx& cx, const uint32_t* in, uint32_t* out, uint32_t N, uint32_t* tmp) {
    const uint32_t tiles = div_up(N, kScanTile);
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, cx.stream, in, out, nullptr, N);
        return;
    }
    uint32_t* sums = tmp;
    uint32_t* sums_scanned = tmp + tiles;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, cx.stream, in, out, sums, N);
    scan_exclusive(cx, sums, sums_scanned, tiles, tmp + 2 * (size_t)tiles);
    hipLaunchKernelGGL(k_scan_add, dim3(div_up(N, 256)), dim3(256), 0, cx.stream, out, sums_scanned, N);
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

constexpr uint32_t kChunk = 32;     // sorted entries per accumulate lane
constexpr uint32_t kSegment = 8;    // buckets per reduction lane

// Runs the whole device pipeline for points/scalars already on the device
// and writes the window sums to host memory (xyzz6 per window).
void run_device(Ctx& cx, const Plan& pl, const void* d_points, const void* d_scalars,
                std::vector<hostec::xyzz6>& wins) {
    const uint32_t nkeys = pl.nwin * pl.B;
    const uint64_t nent = (uint64_t)pl.nwin * pl.n;
    if (nent >= 0x7fffffffULL || (uint64_t)nkeys >= 0x7fffffffULL)
        throw HipFail{"MSM too large for one launch"};
    const uint32_t nchunks = div_up(nent, kChunk);
    const uint32_t nslots = 2 * nchunks;
    const uint32_t L = std::min<uint32_t>(kSegment, pl.B);
    const uint32_t P = pl.B / L;

    cx.ent.ensure(nent * 4);
    cx.rank.ensure(nent * 4);
    cx.skey.ensure(nent * 4);
    cx.sval.ensure(nent * 4);
    cx.hist.ensure(((size_t)nkeys + 1) * 4);
    cx.off.ensure(((size_t)nkeys + 1) * 4);
    cx.scan_tmp.ensure(scan_tmp_words(nkeys + 1) * 4);
    cx.buckets.ensure((size_t)nkeys * sizeof(xyzz));
    cx.slot_key.ensure((size_t)nslots * 4);
    cx.slots.ensure((size_t)nslots * sizeof(xyzz));
    cx.H.ensure((size_t)pl.nwin * P * sizeof(xyzz));

    win_params wp;
    wp.n = pl.n;
    wp.nbytes = pl.nbytes;
    wp.nbits = pl.nbits;
    wp.c = pl.c;
    wp.nwin = pl.nwin;
    wp.bit_base = pl.bit_base;
    wp.B = pl.B;
    wp.vec_scalars = (pl.nbytes == 32 && ((uintptr_t)d_scalars % 16) == 0) ? 1u : 0u;

    hipStream_t st = cx.stream;
    uint32_t* hist = cx.hist.as<uint32_t>();
    uint32_t* off = cx.off.as<uint32_t>();

    HIPCK(hipEventRecord(cx.ev[0], st));
    HIPCK(hipMemsetAsync(hist, 0, ((size_t)nkeys + 1) * 4, st));
    hipLaunchKernelGGL(k_digits, dim3(div_up(pl.n, 256)), dim3(256), 0, st, wp,
                       (const uint8_t*)d_scalars, (const uint4*)d_points, cx.ent.as<uint32_t>(),
                       cx.rank.as<uint32_t>(), hist);
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(cx.ev[1], st));

    scan_exclusive(cx, hist, off, nkeys + 1, cx.scan_tmp.as<uint32_t>());
    hipLaunchKernelGGL(k_scatter, dim3(div_up(nent, 256)), dim3(256), 0, st, wp,
                       cx.ent.as<const uint32_t>(), cx.rank.as<const uint32_t>(), off,
                       cx.skey.as<uint32_t>(), cx.sval.as<uint32_t>());
    HIPCK(hipGetLastError());
    HIPCK(hipMemsetAsync(cx.buckets.p, 0, (size_t)nkeys * sizeof(xyzz), st));
    HIPCK(hipEventRecord(cx.ev[2], st));
    hipLaunchKernelGGL(k_accumulate, dim3(div_up(nchunks, 256)), dim3(256), 0, st,
                       cx.skey.as<const uint32_t>(), cx.sval.as<const uint32_t>(), off + nkeys,
                       (const uint4*)d_points, cx.buckets.as<xyzz>(), cx.slot_key.as<uint32_t>(),
                       cx.slots.as<xyzz>(), kChunk, nchunks);
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(cx.ev[3], st));

    // a bucket holds at most n entries of a window, so it spans at most
    // ceil(n / T) + 1 chunks
    const uint32_t maxspan = div_up(pl.n, kChunk) + 1;
    for (uint32_t d = 1; d < maxspan; d <<= 1)
        hipLaunchKernelGGL(k_partial_round, dim3(div_up(nslots, 256)), dim3(256), 0, st,
                           cx.slot_key.as<const uint32_t>(), cx.slots.as<xyzz>(), off, hist,
                           kChunk, nslots, d);
    hipLaunchKernelGGL(k_partial_final, dim3(div_up(nslots, 256)), dim3(256), 0, st,
                       cx.slot_key.as<const uint32_t>(), cx.slots.as<const xyzz>(),
                       cx.buckets.as<xyzz>(), nslots);
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(cx.ev[4], st));

    hipLaunchKernelGGL(k_reduce_segments, dim3(div_up((uint64_t)pl.nwin * P, 256)), dim3(256), 0, st,
                       cx.buckets.as<const xyzz>(), cx.H.as<xyzz>(), pl.B, L, pl.nwin);
    for (uint32_t step = 1; step < P; step <<= 1)
        hipLaunchKernelGGL(k_pair_sum, dim3(div_up((uint64_t)pl.nwin * (P / (2 * step)), 256)), dim3(256),
                           0, st, cx.H.as<xyzz>(), P, pl.nwin, step);
    HIPCK(hipGetLastError());
    wins.resize(pl.nwin);
    HIPCK(hipMemcpy2DAsync(wins.data(), sizeof(xyzz), cx.H.p, (size_t)P * sizeof(xyzz), sizeof(xyzz),
                           pl.nwin, hipMemcpyDeviceToHost, st));
    HIPCK(hipEventRecord(cx.ev[5], st));
    HIPCK(hipStreamSynchronize(st));
    float ms;
    for (int k = 0; k < 5; k++) {
        HIPCK(hipEventElapsedTime(&ms, cx.ev[k], cx.ev[k + 1]));
        g_phase[k] = ms;
    }
}

void finish(const Plan& pl, const std::vector<hostec::xyzz6>& wins, blst_p1* ret) {
    auto t0 = std::chrono::steady_clock::now();
    hostec::xyzz6 r;
    if (pl.fold) hostec::fold_windows(r, wins.data(), (int)pl.nwin, (int)pl.c);
    else r = wins[0];
    hostec::jac6 j;
    hostec::to_jac(j, r);
    static_assert(sizeof(hostec::jac6) == sizeof(blst_p1), "blst_p1 layout");
    std::memcpy(ret, &j, sizeof(j));
    auto t1 = std::chrono::steady_clock::now();
    g_phase[5] = std::chrono::duration<double, std::milli>(t1 - t0).count();
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
        const uint8_t* cur;
        if (!ended && items[i]) cur = reinterpret_cast<const uint8_t*>(items[i]);
        else { ended = true; cur = prev + stride; }
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
        cx.pts.ensure(npoints * sizeof(blst_p1_affine));
        cx.scl.ensure(npoints * nbytes + 16);
        HIPCK(hipMemcpyAsync(cx.pts.p, hp, npoints * sizeof(blst_p1_affine), hipMemcpyHostToDevice, cx.stream));
        HIPCK(hipMemcpyAsync(cx.scl.p, hs, npoints * nbytes, hipMemcpyHostToDevice, cx.stream));
        Plan pl;
        pl.n = (uint32_t)npoints;
        pl.nbits = (uint32_t)nbits;
        pl.nbytes = (uint32_t)nbytes;
        pl.c = (uint32_t)(window ? window : device_window(npoints, nbits));
        pl.nwin = fold ? (uint32_t)(1 + nbits / pl.c) : 1u;
        pl.B = 1u << (pl.c - 1);
        pl.bit_base = (uint32_t)bit0;
        pl.fold = fold;
        std::vector<hostec::xyzz6> wins;
        run_device(cx, pl, cx.pts.p, cx.scl.p, wins);
        finish(pl, wins, ret);
    } catch (const HipFail& f) {
        g_err = f.msg;
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    } catch (const std::exception& e) {
        g_err = e.what();
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_phase[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    g_err.clear();
    return 0;
}

// Comb table T[m][j-1] = [j * 2^(8m)] G, m < 32, j < 256, affine, on the
// device; built once per context on the host (8160 additions) with one
// batched inversion.
void ensure_gtable(Ctx& cx) {
    if (cx.gtable_ready) return;
    static const uint64_t GX[6] = {0x5cb38790fd530c16ULL, 0x7817fc679976fff5ULL, 0x154f95c7143ba1c1ULL,
                                   0xf0ae6acdf3d0e747ULL, 0xedce6ecc21dbf440ULL, 0x120177419e0bfb75ULL};
    static const uint64_t GY[6] = {0xbaac93d50ce72271ULL, 0x8c22631a7918fd8eULL, 0xdd595f13570725ceULL,
                                   0x51ac582950405194ULL, 0x0e1c8c3fad0059c0ULL, 0x0bbc3efc5008a26aULL};
    hostec::xyzz6 base;
    std::memcpy(base.x.l, GX, 48);
    std::memcpy(base.y.l, GY, 48);
    base.zz = hostec::one();
    base.zzz = hostec::one();
    std::vector<hostec::xyzz6> T(32 * 255);
    for (int m = 0; m < 32; m++) {
        hostec::xyzz6 acc = base;
        for (int j = 1; j < 256; j++) {
            if (j > 1) hostec::addp(acc, acc, base);
            T[m * 255 + (j - 1)] = acc;
        }
        for (int k = 0; k < 8; k++) hostec::dbl(base, base);
    }
    std::vector<hostec::aff6> A(T.size());
    hostec::batch_to_affine(A.data(), T.data(), T.size());
    cx.gtable.ensure(A.size() * sizeof(hostec::aff6));
    HIPCK(hipMemcpy(cx.gtable.p, A.data(), A.size() * sizeof(hostec::aff6), hipMemcpyHostToDevice));
    cx.gtable_ready = true;
}

bool heads_ok(const void* const* xs, const void* const* ys, size_t n) {
    return n == 0 || (xs != nullptr && ys != nullptr && xs[0] != nullptr && ys[0] != nullptr);
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
    auto t0 = std::chrono::steady_clock::now();
    std::memset(ret, 0, sizeof(*ret));
    if (npoints == 0 || nbits == 0) { g_err.clear(); return 0; }
    if (nbits > 256 || window > 24 || !device_points || !device_scalars ||
        ((uintptr_t)device_points % 16) != 0) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        Plan pl;
        pl.n = (uint32_t)npoints;
        pl.nbits = (uint32_t)nbits;
        pl.nbytes = (uint32_t)((nbits + 7) / 8);
        pl.c = (uint32_t)(window ? window : device_window(npoints, nbits));
        pl.nwin = (uint32_t)(1 + nbits / pl.c);
        pl.B = 1u << (pl.c - 1);
        pl.bit_base = 0;
        pl.fold = true;
        std::vector<hostec::xyzz6> wins;
        run_device(cx, pl, device_points, device_scalars, wins);
        finish(pl, wins, ret);
    } catch (const HipFail& f) {
        g_err = f.msg;
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_phase[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    g_err.clear();
    return 0;
}

int blst_hip_booth_digits(int32_t* out, const blst_byte* scalars, size_t npoints, size_t nbits,
                          size_t bit0, size_t window) {
    if (npoints == 0) return 0;
    if (!out || !scalars || nbits == 0 || nbits > 256 || window < 1 || window > 24 || bit0 > nbits) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        const size_t nbytes = (nbits + 7) / 8;
        cx.scl.ensure(npoints * nbytes + 16);
        cx.dbg.ensure(npoints * 4);
        HIPCK(hipMemcpyAsync(cx.scl.p, scalars, npoints * nbytes, hipMemcpyHostToDevice, cx.stream));
        hipLaunchKernelGGL(k_digits_debug, dim3(div_up(npoints, 256)), dim3(256), 0, cx.stream,
                           cx.scl.as<const uint8_t>(), (uint32_t)npoints, (uint32_t)nbytes, (uint32_t)nbits,
                           (uint32_t)bit0, (uint32_t)window, cx.dbg.as<int32_t>());
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, cx.dbg.p, npoints * 4, hipMemcpyDeviceToHost, cx.stream));
        HIPCK(hipStreamSynchronize(cx.stream));
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    }
    g_err.clear();
    return 0;
}

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

void blst_hip_device_free(void* device_ptr) {
    if (device_ptr) (void)hipFree(device_ptr);
}

int blst_hip_p1s_mult_generator(blst_p1_affine* out, const blst_byte* scalars, size_t npoints,
                                size_t nbits) {
    if (npoints == 0) { g_err.clear(); return 0; }
    if (!out || !scalars || nbits == 0 || nbits > 256 || npoints >= 0x7fffffffULL) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        ensure_gtable(cx);
        const size_t nbytes = (nbits + 7) / 8;
        const uint32_t Q = 64;
        cx.scl.ensure(npoints * nbytes + 16);
        cx.gen_tmp.ensure(npoints * sizeof(xyzz));
        cx.gen_out.ensure(npoints * sizeof(aff));
        HIPCK(hipMemcpyAsync(cx.scl.p, scalars, npoints * nbytes, hipMemcpyHostToDevice, cx.stream));
        hipLaunchKernelGGL(k_fixed_base, dim3(div_up(npoints, 256)), dim3(256), 0, cx.stream,
                           cx.scl.as<const uint8_t>(), (uint32_t)npoints, (uint32_t)nbytes, (uint32_t)nbits,
                           cx.gtable.as<const uint4>(), cx.gen_tmp.as<xyzz>());
        hipLaunchKernelGGL(k_to_affine, dim3(div_up(div_up(npoints, Q), 256)), dim3(256), 0, cx.stream,
                           cx.gen_tmp.as<const xyzz>(), cx.gen_out.as<aff>(), (uint32_t)npoints, Q);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, cx.gen_out.p, npoints * sizeof(aff), hipMemcpyDeviceToHost, cx.stream));
        HIPCK(hipStreamSynchronize(cx.stream));
    } catch (const HipFail& f) {
        g_err = f.msg;
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

void blst_hip_last_phase_ms(double out[7]) {
    for (int k = 0; k < 7; k++) out[k] = g_phase[k];
}

}  // extern "C"
