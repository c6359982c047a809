// Test and workload-generation hooks, built as libblst_hip_testkit.so beside
// the product library (which it links against) and never part of the drop-in
// surface:
//   * blst_hip_fp_ops: one device field operation (fp.hip.h) applied over
//     arrays, so the [0, 2p) redundant field layer can be compared with exact
//     big-integer arithmetic on chosen inputs (p, 2p - 1, R mod p, ...);
//   * blst_hip_fq_ops: the same for the 28-bit limb field of the G1 bucket
//     accumulation (fq.hip.h), including its lazy values and its
//     conversions from and to the 32-bit word image;
//   * blst_hip_booth_digits: the signed window digits as the sort kernels
//     read them;
//   * blst_hip_point_ops_check: the four-lane point formulas against the
//     one-lane ones;
//   * blst_hip_p1s_mult_generator: [k_i]G on the device, for synthetic bases
//     (the seeded scalar generators are in synth.cpp).
#include "pipeline.h"
#include "ec_quad.hip.h"
#include "fq.hip.h"
#include "../../include/blst_hip_testkit.h"

#include <cstring>
#include <mutex>
#include <vector>

namespace bls {

// field operations of the device layer, selected by op
enum : uint32_t {
    kOpMul = 0, kOpSqr, kOpMulSum, kOpAdd, kOpSub, kOpDbl, kOpNeg, kOpCneg, kOpCanon, kOpIsZero, kOpInv,
    kOpCount
};

__device__ __forceinline__ void load_fp(fp& r, const uint4* __restrict__ src, uint32_t i) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 v = src[3 * i + k];
        r.l[4 * k] = v.x;
        r.l[4 * k + 1] = v.y;
        r.l[4 * k + 2] = v.z;
        r.l[4 * k + 3] = v.w;
    }
}

__device__ __forceinline__ void store_fp(uint4* __restrict__ dst, uint32_t i, const fp& r) {
#pragma unroll
    for (int k = 0; k < 3; k++) dst[3 * i + k] = make_uint4(r.l[4 * k], r.l[4 * k + 1], r.l[4 * k + 2], r.l[4 * k + 3]);
}

__global__ void __launch_bounds__(256)
k_fp_ops(uint32_t op, uint4* __restrict__ out, const uint4* __restrict__ A, const uint4* __restrict__ B,
         const uint4* __restrict__ C, const uint4* __restrict__ D, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fp a, b, c, d, r;
    load_fp(a, A, i);
    load_fp(b, B, i);
    switch (op) {
    case kOpMul: fp_mul(r, a, b); break;
    case kOpSqr: fp_sqr(r, a); break;
    case kOpMulSum:
        load_fp(c, C, i);
        load_fp(d, D, i);
        fp_mul_sum(r, a, b, c, d);
        break;
    case kOpAdd: fp_add(r, a, b); break;
    case kOpSub: fp_sub(r, a, b); break;
    case kOpDbl: fp_dbl(r, a); break;
    case kOpNeg: fp_neg(r, a); break;
    case kOpCneg: fp_cneg(r, a, (b.l[0] & 1) != 0); break;
    case kOpCanon: fp_canon(r, a); break;
    case kOpIsZero:
        fp_zero(r);
        r.l[0] = fp_is_zero(a) ? 1u : 0u;
        break;
    default: fp_inv(r, a); break;
    }
    store_fp(out, i, r);
}

// operations of the 28-bit limb field, selected by op; rows of 14 words
enum : uint32_t {
    kFqMul = 0, kFqSqr, kFqMulSum, kFqReduce, kFqSubLazy, kFqNegLazy, kFqX3, kFqIsZero, kFqFromWords8, kFqFromWords4,
    kFqFromWords0, kFqToWords, kFqTimes4, kFqReduceWide, kFqCanon, kFqCount
};

__global__ void __launch_bounds__(256)
k_fq_ops(uint32_t op, uint32_t* __restrict__ out, const uint32_t* __restrict__ A, const uint32_t* __restrict__ B,
         const uint32_t* __restrict__ C, const uint32_t* __restrict__ D, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fq a, b, c, d, r;
#pragma unroll
    for (int k = 0; k < 14; k++) {
        a.l[k] = A[14 * i + k];
        b.l[k] = B[14 * i + k];
        c.l[k] = C[14 * i + k];
        d.l[k] = D[14 * i + k];
    }
    fq_zero(r);
    switch (op) {
    case kFqMul: fq_mul(r, a, b); break;
    case kFqSqr: fq_sqr(r, a); break;
    case kFqMulSum: fq_mul_sum(r, a, b, c, d); break;
    case kFqReduce: fq_reduce(r, a); break;
    case kFqSubLazy: fq_sub_lazy(r, a, b); break;
    case kFqNegLazy: fq_neg_lazy(r, a); break;
    case kFqX3: fq_x3(r, a, b, c, (d.l[0] & 1) != 0); break;
    case kFqIsZero: r.l[0] = fq_is_zero(a) ? 1u : 0u; break;
    case kFqFromWords8: fq_from_words<8>(r, a.l); break;
    case kFqFromWords4: fq_from_words<4>(r, a.l); break;
    case kFqFromWords0: fq_from_words<0>(r, a.l); break;
    case kFqToWords: fq_to_words(r.l, a); break;
    case kFqTimes4: fq_times4(r, a); break;
    case kFqReduceWide: fq_reduce_wide(r, a); break;
    default: fq_canon(r, a); break;
    }
#pragma unroll
    for (int k = 0; k < 14; k++) out[14 * i + k] = r.l[k];
}

__global__ void k_digits_debug(const uint8_t* __restrict__ scalars, uint32_t n, uint32_t nbytes, uint32_t nbits,
                               uint32_t bit0, uint32_t c, int32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the tiled sort's reader (word reads for aligned 32-byte scalars)
    const bool vec = nbytes == 32 && ((uintptr_t)scalars & 15) == 0;
    out[i] = booth_digit_at(scalars + (size_t)i * nbytes, vec, bit0, c, nbits);
}

// Issue-rate probe of v_mad_u64_u32 (the multiply-add of every Montgomery
// product): eight independent 64-bit accumulators per lane, so a wave always
// has a multiply-add whose inputs are ready, 64 rounds per pass.  The sum of
// the accumulators is stored (a vector store) so nothing is optimised away.
__global__ void __launch_bounds__(256)
k_mad_issue(uint64_t* __restrict__ out, uint32_t passes, uint32_t seed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t x = seed ^ (i * 0x9e3779b9u), y = x * 0x85ebca6bu + 1u;
    uint64_t a0 = i, a1 = i + 1, a2 = i + 2, a3 = i + 3, a4 = i + 4, a5 = i + 5, a6 = i + 6, a7 = i + 7;
#pragma unroll 1
    for (uint32_t p = 0; p < passes; p++) {
#pragma unroll
        for (int r = 0; r < 64; r++) {
            asm volatile(
                "v_mad_u64_u32 %0, vcc, %8, %9, %0\n\t"
                "v_mad_u64_u32 %1, vcc, %8, %9, %1\n\t"
                "v_mad_u64_u32 %2, vcc, %8, %9, %2\n\t"
                "v_mad_u64_u32 %3, vcc, %8, %9, %3\n\t"
                "v_mad_u64_u32 %4, vcc, %8, %9, %4\n\t"
                "v_mad_u64_u32 %5, vcc, %8, %9, %5\n\t"
                "v_mad_u64_u32 %6, vcc, %8, %9, %6\n\t"
                "v_mad_u64_u32 %7, vcc, %8, %9, %7"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
                : "v"(x), "v"(y)
                : "vcc");
        }
        x += 1;
    }
    out[i] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

void launch_ba_probe(hipStream_t stream, int K, bool batch_affine, const uint4* pts, const uint32_t* idx, fp* pi,
                     void* out, uint32_t npairs);
void launch_gather_probe(hipStream_t stream, int stride_bytes, const uint4* pts, const uint32_t* idx, void* out,
                         uint32_t nentries);

namespace host {
namespace {

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Device buffers of the hooks, one set per process.  The hooks work on the
// null stream and create none of their own: a process gets few hardware
// queues (four by default), and a stream created here before the MSM
// library's streams would make two of those share one queue, which runs
// their work one after the other.
struct Kit {
    std::mutex mu;
    hipStream_t stream = nullptr;   // the null stream
    DevBuf in[4], out, scl, gtable, gen_tmp, gen_out;
    bool gtable_ready = false;
    void init() {}
};

Kit& kit() {
    static Kit* k = new Kit();   // never destroyed: the runtime may be gone at exit
    return *k;
}

// Comb table T[m][j-1] = [j * 2^(8m)] G, m < 32, j < 256, affine, built once
// on the host (8160 additions) with one batched inversion.
void ensure_gtable(Kit& k) {
    if (k.gtable_ready) return;
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
        for (int s = 0; s < 8; s++) hostec::dbl(base, base);
    }
    std::vector<hostec::aff6> A(T.size());
    hostec::batch_to_affine(A.data(), T.data(), T.size());
    k.gtable.ensure(A.size() * sizeof(hostec::aff6));
    HIPCK(hipMemcpy(k.gtable.p, A.data(), A.size() * sizeof(hostec::aff6), hipMemcpyHostToDevice));
    k.gtable_ready = true;
}

// Lanes of the XYZZ -> affine conversion: enough to fill the device with one
// inversion per lane.
constexpr size_t kGenAffineLanes = (size_t)1 << 17;

}  // namespace
}  // namespace host
}  // namespace bls

using namespace bls;
using namespace bls::host;

extern "C" {

void blst_hip_set_alloc_cap(size_t bytes) { g_alloc_cap.store(bytes); }

void blst_hip_set_stage_failure(int range) { g_stage_fail_range.store(range); }

int blst_hip_fp_ops(int op, void* out, const void* a, const void* b, const void* c, const void* d, size_t n) {
    if (n == 0) { g_err.clear(); return 0; }
    if (op < 0 || op >= (int)kOpCount || !out || !a || !b || n >= (1u << 26) ||
        (op == (int)kOpMulSum && (!c || !d))) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        k.init();
        const size_t bytes = n * 48;
        const void* src[4] = {a, b, c, d};
        const int nin = op == (int)kOpMulSum ? 4 : 2;
        for (int j = 0; j < 4; j++) k.in[j].ensure(bytes);
        for (int j = 0; j < nin; j++)
            HIPCK(hipMemcpyAsync(k.in[j].p, src[j], bytes, hipMemcpyHostToDevice, k.stream));
        k.out.ensure(bytes);
        hipLaunchKernelGGL(k_fp_ops, dim3(div_up(n, 256)), dim3(256), 0, k.stream, (uint32_t)op, k.out.as<uint4>(),
                           k.in[0].as<const uint4>(), k.in[1].as<const uint4>(), k.in[2].as<const uint4>(),
                           k.in[3].as<const uint4>(), (uint32_t)n);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, k.out.p, bytes, hipMemcpyDeviceToHost, k.stream));
        HIPCK(hipStreamSynchronize(k.stream));
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_fq_ops(int op, uint32_t* out, const uint32_t* a, const uint32_t* b, const uint32_t* c,
                    const uint32_t* d, size_t n) {
    if (n == 0) { g_err.clear(); return 0; }
    if (op < 0 || op >= (int)kFqCount || !out || !a || !b || !c || !d || n >= (1u << 26)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        const size_t bytes = n * 56;
        const void* src[4] = {a, b, c, d};
        for (int j = 0; j < 4; j++) {
            k.in[j].ensure(bytes);
            HIPCK(hipMemcpyAsync(k.in[j].p, src[j], bytes, hipMemcpyHostToDevice, k.stream));
        }
        k.out.ensure(bytes);
        hipLaunchKernelGGL(k_fq_ops, dim3(div_up(n, 256)), dim3(256), 0, k.stream, (uint32_t)op,
                           k.out.as<uint32_t>(), k.in[0].as<const uint32_t>(), k.in[1].as<const uint32_t>(),
                           k.in[2].as<const uint32_t>(), k.in[3].as<const uint32_t>(), (uint32_t)n);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, k.out.p, bytes, hipMemcpyDeviceToHost, k.stream));
        HIPCK(hipStreamSynchronize(k.stream));
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_mad_issue_rate(double* mad_per_s) {
    if (!mad_per_s) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        int dev = 0, cus = 0;
        HIPCK(hipGetDevice(&dev));
        HIPCK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        // eight waves per SIMD on every CU
        const uint32_t blocks = (uint32_t)cus * 8, threads = 256, passes = 512;
        k.out.ensure((size_t)blocks * threads * 8);
        hipEvent_t e0, e1;
        HIPCK(hipEventCreate(&e0));
        HIPCK(hipEventCreate(&e1));
        float best = 0;
        for (int rep = 0; rep < 4; rep++) {
            HIPCK(hipEventRecord(e0, k.stream));
            hipLaunchKernelGGL(k_mad_issue, dim3(blocks), dim3(threads), 0, k.stream, k.out.as<uint64_t>(), passes,
                               (uint32_t)rep);
            HIPCK(hipGetLastError());
            HIPCK(hipEventRecord(e1, k.stream));
            HIPCK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && (best == 0 || ms < best)) best = ms;   // the first launch is a warm-up
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const double mads = (double)blocks * threads * passes * 64 * 8;
        *mad_per_s = mads / (best * 1e-3);
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_batch_affine_probe(const void* device_points, const uint32_t* device_index, size_t npairs,
                                int pairs_per_lane, double ns_per_pair[2], void* device_affine_out,
                                void* device_xyzz_out) {
    const int K = pairs_per_lane;
    if (!device_points || !ns_per_pair || !device_affine_out || !device_xyzz_out || (K != 8 && K != 32 && K != 128) ||
        npairs == 0 || npairs % ((size_t)K * 256) != 0 || npairs >= (1u << 28) || ((uintptr_t)device_points & 15)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        k.scl.ensure(npairs * sizeof(fp));   // prefix products
        hipEvent_t e0, e1;
        HIPCK(hipEventCreate(&e0));
        HIPCK(hipEventCreate(&e1));
        for (int mode = 0; mode < 2; mode++) {
            float best = 0;
            for (int rep = 0; rep < 4; rep++) {
                HIPCK(hipEventRecord(e0, k.stream));
                launch_ba_probe(k.stream, K, mode == 0, static_cast<const uint4*>(device_points), device_index,
                                k.scl.as<fp>(), mode == 0 ? device_affine_out : device_xyzz_out, (uint32_t)npairs);
                HIPCK(hipGetLastError());
                HIPCK(hipEventRecord(e1, k.stream));
                HIPCK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCK(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0 && (best == 0 || ms < best)) best = ms;   // the first launch is a warm-up
            }
            ns_per_pair[mode] = best * 1e6 / (double)npairs;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_gather_probe(const void* device_points, size_t stride_bytes, const uint32_t* device_index,
                          size_t nentries, double* ns_per_entry, void* device_xyzz_out) {
    if (!device_points || !device_index || !ns_per_entry || !device_xyzz_out ||
        (stride_bytes != 96 && stride_bytes != 128) || nentries == 0 || nentries % (32 * 256) != 0 ||
        nentries >= (1u << 28) || ((uintptr_t)device_points & 15)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        hipEvent_t e0, e1;
        HIPCK(hipEventCreate(&e0));
        HIPCK(hipEventCreate(&e1));
        float best = 0;
        for (int rep = 0; rep < 4; rep++) {
            HIPCK(hipEventRecord(e0, k.stream));
            launch_gather_probe(k.stream, (int)stride_bytes, static_cast<const uint4*>(device_points), device_index,
                                device_xyzz_out, (uint32_t)nentries);
            HIPCK(hipGetLastError());
            HIPCK(hipEventRecord(e1, k.stream));
            HIPCK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && (best == 0 || ms < best)) best = ms;   // the first launch is a warm-up
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ns_per_entry = best * 1e6 / (double)nentries;
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_point_ops_check(const void* device_points, size_t npoints) {
    if (npoints == 0) return 0;
    if (!device_points || npoints > (1u << 24) || ((uintptr_t)device_points & 15)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        k.init();
        k.out.ensure(64);
        HIPCK(hipMemsetAsync(k.out.p, 0, 4, k.stream));
        launch_k_quad_check(dim3(div_up(4 * npoints, 256)), dim3(256), k.stream,
                            static_cast<const uint4*>(device_points), (uint32_t)npoints, k.out.as<uint32_t>());
        HIPCK(hipGetLastError());
        uint32_t bad = 0;
        HIPCK(hipMemcpyAsync(&bad, k.out.p, 4, hipMemcpyDeviceToHost, k.stream));
        HIPCK(hipStreamSynchronize(k.stream));
        g_err.clear();
        return (int)bad;
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    }
}

int blst_hip_booth_digits(int32_t* out, const blst_byte* scalars, size_t npoints, size_t nbits, size_t bit0,
                          size_t window) {
    if (npoints == 0) return 0;
    if (!out || !scalars || nbits == 0 || nbits > 256 || window < 1 || window > 24 || bit0 > nbits ||
        npoints >= 0x7fffffffULL) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        k.init();
        const size_t nbytes = (nbits + 7) / 8;
        k.scl.ensure(npoints * nbytes + 16);
        k.out.ensure(npoints * 4);
        HIPCK(hipMemcpyAsync(k.scl.p, scalars, npoints * nbytes, hipMemcpyHostToDevice, k.stream));
        hipLaunchKernelGGL(k_digits_debug, dim3(div_up(npoints, 256)), dim3(256), 0, k.stream,
                           k.scl.as<const uint8_t>(), (uint32_t)npoints, (uint32_t)nbytes, (uint32_t)nbits,
                           (uint32_t)bit0, (uint32_t)window, k.out.as<int32_t>());
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, k.out.p, npoints * 4, hipMemcpyDeviceToHost, k.stream));
        HIPCK(hipStreamSynchronize(k.stream));
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_p1s_mult_generator(blst_p1_affine* out, const blst_byte* scalars, size_t npoints, size_t nbits) {
    if (npoints == 0) { g_err.clear(); return 0; }
    if (!out || !scalars || nbits == 0 || nbits > 256 || npoints >= 0x7fffffffULL) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Kit& k = kit();
        std::lock_guard<std::mutex> lk(k.mu);
        k.init();
        ensure_gtable(k);
        const size_t nbytes = (nbits + 7) / 8;
        const uint32_t Q = (uint32_t)std::min<size_t>(64, std::max<size_t>(1, npoints / kGenAffineLanes));
        k.scl.ensure(npoints * nbytes + 16);
        k.gen_tmp.ensure(npoints * sizeof(xyzz));
        k.gen_out.ensure(npoints * sizeof(aff));
        HIPCK(hipMemcpyAsync(k.scl.p, scalars, npoints * nbytes, hipMemcpyHostToDevice, k.stream));
        launch_k_fixed_base(dim3(div_up(npoints, 256)), dim3(256), k.stream, k.scl.as<const uint8_t>(),
                            (uint32_t)npoints, (uint32_t)nbytes, (uint32_t)nbits, k.gtable.as<const uint4>(),
                            k.gen_tmp.as<xyzz>());
        launch_k_to_affine(dim3(div_up(div_up(npoints, Q), 256)), dim3(256), k.stream, k.gen_tmp.as<const xyzz>(),
                           k.gen_out.as<aff>(), (uint32_t)npoints, Q);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, k.gen_out.p, npoints * sizeof(aff), hipMemcpyDeviceToHost, k.stream));
        HIPCK(hipStreamSynchronize(k.stream));
    } catch (const HipFail& f) {
        g_err = f.msg;
        return -1;
    }
    g_err.clear();
    return 0;
}

}  // extern "C"
