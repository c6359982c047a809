// Bucket accumulation: chunks of the sorted list per lane, XYZZ mixed
// additions of gathered affine points.
#include "msm_common.hip.h"
#include "fq.hip.h"
#include "launch.h"

namespace bls {

// The running bucket sum of one lane.  G2 sums in the XYZZ form of
// ec.hip.h.  G1 sums in fq (28-bit limbs, no carry folds in its products)
// and converts where a bucket sum is loaded or stored: the stored sums keep
// the 32-bit layout that the merge and reduction kernels read.
template <class F>
struct bucket_sum {
    xyzz_t<F> a;
    __device__ __forceinline__ void set_inf() { xyzz_set_inf(a); }
    __device__ __forceinline__ void load(const xyzz_t<F>* __restrict__ src) { load_xyzz(a, src); }
    __device__ __forceinline__ void store(xyzz_t<F>* __restrict__ dst) const { store_xyzz(dst, a); }
    // point v & 0x7fffffff of pts, negated when bit 31 is set
    __device__ __forceinline__ void add(const uint4* __restrict__ pts, uint32_t v) {
        aff_t<F> q;
        load_aff(q, pts, v & 0x7fffffffu);
        // an all-zero affine point is infinity and contributes nothing
        // (Pointxyzz::dadd with an all-zero affine input, point.cpp:2036-2039)
        if (!aff_is_inf(q)) {
            if (v >> 31) fp_neg(q.y, q.y);
            xyzz_madd(a, q);
        }
    }
};

template <>
struct bucket_sum<fp> {
    fq_xyzz a;
    __device__ __forceinline__ void set_inf() { fq_xyzz_set_inf(a); }
    __device__ __forceinline__ void load(const xyzz* __restrict__ src) {
        xyzz w;
        load_xyzz(w, src);
        fq_xyzz_from_words(a, w.x.l, w.y.l, w.zz.l, w.zzz.l);
    }
    __device__ __forceinline__ void store(xyzz* __restrict__ dst) const {
        xyzz w;
        fq_xyzz_to_words(w.x.l, w.y.l, w.zz.l, w.zzz.l, a);
        store_xyzz(dst, w);
    }
    __device__ __forceinline__ void add(const uint4* __restrict__ pts, uint32_t v) {
        aff q;
        load_aff(q, pts, v & 0x7fffffffu);
        if (!aff_is_inf(q)) {
            if (v >> 31) fp_neg(q.y, q.y);   // 2p - y: below 2p, as fq_xyzz_madd takes it
            fq x, y16;
            fq_from_words<8>(x, q.x.l);
            fq_from_words<4>(y16, q.y.l);
            fq_xyzz_madd(a, x, y16);
        }
    }
};

// One lane per chunk of T sorted entries.  A bucket segment that starts and
// ends inside the chunk is final and goes to buckets[key].  Otherwise, with
// k0 the chunk holding the bucket's first entry: the segment in chunk k0 goes
// to slot 2*k0+1, the segment in any later chunk k to slot 2*k.  The lane of
// chunk k0 also lists a bucket of more than two pieces for the merge (its
// entry count cnt[key] from offset off[key] fixes how many chunks it spans).
// With prev (the entry counts of earlier point ranges of the same MSM), a
// bucket that already holds a sum continues from it: the lane holding the
// bucket's first entry starts from buckets[key] instead of infinity.
// waves per SIMD the kernel is compiled for: three for G1 (at most 168
// VGPRs), what the register use allows for G2
template <class F> constexpr int kAccumulateBlocksPerCU = 1;
template <> constexpr int kAccumulateBlocksPerCU<fp> = 3;

template <class F>
__global__ void __launch_bounds__(256, kAccumulateBlocksPerCU<F>)
k_accumulate(const uint2* __restrict__ sorted,
             const uint32_t* __restrict__ total, const job_t* __restrict__ jobs,
             const uint32_t* __restrict__ kbase, uint32_t njobs, xyzz_t<F>* __restrict__ buckets,
             uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots, uint32_t T, uint32_t nchunks,
             const uint32_t* __restrict__ prev, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
             merge_lists_t lists) {
    // Each lane's next eight sorted entries (one 64-byte segment), loaded at
    // once and read from LDS: read one by one from memory, a wave's 64 lanes
    // touch 64 lines T entries apart, and the point gathers in between evict
    // those lines before the next entry is read.  Slot q of lane l is
    // stage[q][l], so the writes and reads of a wave fall on distinct banks.
    __shared__ uint4 stage[4][256];
    const bool staged = (T & 7) == 0;
    const uint32_t chunk = blockIdx.x * blockDim.x + threadIdx.x;
    if (chunk >= nchunks) return;
    slot_key[2 * chunk] = kInvalid;
    slot_key[2 * chunk + 1] = kInvalid;
    const uint32_t M = *total;
    const uint32_t start = chunk * T;
    if (start >= M) return;
    const uint32_t end = min(start + T, M);
    uint32_t cur = sorted[start].x;
    uint32_t j = find_owner(kbase, njobs, cur);
    uint32_t next_kb = (j + 1 < njobs) ? kbase[j + 1] : 0xffffffffu;
    const uint4* pts = jobs[j].points;
    const bool before = start > 0 && sorted[start - 1].x == cur;
    bool first = true;
    bucket_sum<F> acc;
    if (prev && !before && prev[cur]) acc.load(&buckets[cur]);
    else acc.set_inf();
    #pragma unroll 1
    for (uint32_t e = start; e < end; e++) {
        uint2 kv;
        if (staged) {
            const uint32_t i = (e - start) & 7;
            if (i == 0) {
                // start and e are multiples of 8: 64-byte aligned
                const uint4* src = reinterpret_cast<const uint4*>(sorted + e);
#pragma unroll
                for (uint32_t q = 0; q < 4; q++) {
                    if (e + 2 * q + 1 < end) {
                        stage[q][threadIdx.x] = src[q];
                    } else if (e + 2 * q < end) {
                        const uint2 t = sorted[e + 2 * q];
                        stage[q][threadIdx.x] = make_uint4(t.x, t.y, 0u, 0u);
                    }
                }
            }
            const uint4 v = stage[i >> 1][threadIdx.x];
            kv = (i & 1) ? make_uint2(v.z, v.w) : make_uint2(v.x, v.y);
        } else {
            kv = sorted[e];
        }
        const uint32_t key = kv.x;
        if (key != cur) {
            if (first && before) {
                slot_key[2 * chunk] = cur;
                acc.store(&slots[2 * chunk]);
            } else {
                acc.store(&buckets[cur]);
            }
            first = false;
            cur = key;
            if (prev && prev[key]) acc.load(&buckets[key]);
            else acc.set_inf();
            if (key >= next_kb) {
                while (j + 1 < njobs && kbase[j + 1] <= key) j++;
                next_kb = (j + 1 < njobs) ? kbase[j + 1] : 0xffffffffu;
                pts = jobs[j].points;
            }
        }
        acc.add(pts, kv.y);
    }
    const bool after = end < M && sorted[end].x == cur;
    if (first && before) {
        slot_key[2 * chunk] = cur;
        acc.store(&slots[2 * chunk]);
    } else if (after) {
        slot_key[2 * chunk + 1] = cur;
        acc.store(&slots[2 * chunk + 1]);
        const uint32_t m = (off[cur] + cnt[cur] - 1) / T - chunk + 1;   // pieces
        if (m > 2) list_bucket(lists, m, cur);
    } else {
        acc.store(&buckets[cur]);
    }
}

template <class F>
void launch_k_accumulate(dim3 grid, dim3 block, hipStream_t stream, const uint2* __restrict__ sorted, const uint32_t* __restrict__ total, const job_t* __restrict__ jobs, const uint32_t* __restrict__ kbase, uint32_t njobs, xyzz_t<F>* __restrict__ buckets, uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots, uint32_t T, uint32_t nchunks, const uint32_t* __restrict__ prev, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, merge_lists_t lists) {
    hipLaunchKernelGGL(k_accumulate<F>, grid, block, 0, stream, sorted, total, jobs, kbase, njobs, buckets, slot_key, slots, T, nchunks, prev, off, cnt, lists);
}

template void launch_k_accumulate<fp>(dim3, dim3, hipStream_t, const uint2*, const uint32_t*, const job_t*, const uint32_t*, uint32_t, xyzz*, uint32_t*, xyzz*, uint32_t, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, merge_lists_t);
template void launch_k_accumulate<fp2>(dim3, dim3, hipStream_t, const uint2*, const uint32_t*, const job_t*, const uint32_t*, uint32_t, xyzz2*, uint32_t*, xyzz2*, uint32_t, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, merge_lists_t);

}  // namespace bls
