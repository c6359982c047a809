// Bucket accumulation: chunks of the sorted list per lane, XYZZ mixed
// additions of gathered affine points.
#include "msm_common.hip.h"
#include "launch.h"

namespace bls {

// One lane per chunk of T sorted entries.  A bucket segment that starts and
// ends inside the chunk is final and goes to buckets[key].  Otherwise, with
// k0 the chunk holding the bucket's first entry: the segment in chunk k0 goes
// to slot 2*k0+1, the segment in any later chunk k to slot 2*k.  The lane of
// chunk k0 also lists a bucket of more than two pieces for the merge (its
// entry count cnt[key] from offset off[key] fixes how many chunks it spans).
// With prev (the entry counts of earlier point ranges of the same MSM), a
// bucket that already holds a sum continues from it: the lane holding the
// bucket's first entry starts from buckets[key] instead of infinity.
template <class F>
__global__ void __launch_bounds__(256)
k_accumulate(const uint2* __restrict__ sorted,
             const uint32_t* __restrict__ total, const job_t* __restrict__ jobs,
             const uint32_t* __restrict__ kbase, uint32_t njobs, xyzz_t<F>* __restrict__ buckets,
             uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots, uint32_t T, uint32_t nchunks,
             const uint32_t* __restrict__ prev, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
             merge_lists_t lists) {
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
    xyzz_t<F> acc;
    if (prev && !before && prev[cur]) load_xyzz(acc, &buckets[cur]);
    else xyzz_set_inf(acc);
    #pragma unroll 1
    for (uint32_t e = start; e < end; e++) {
        const uint2 kv = sorted[e];
        const uint32_t key = kv.x;
        if (key != cur) {
            if (first && before) {
                slot_key[2 * chunk] = cur;
                store_xyzz(&slots[2 * chunk], acc);
            } else {
                store_xyzz(&buckets[cur], acc);
            }
            first = false;
            cur = key;
            if (prev && prev[key]) load_xyzz(acc, &buckets[key]);
            else xyzz_set_inf(acc);
            if (key >= next_kb) {
                while (j + 1 < njobs && kbase[j + 1] <= key) j++;
                next_kb = (j + 1 < njobs) ? kbase[j + 1] : 0xffffffffu;
                pts = jobs[j].points;
            }
        }
        const uint32_t v = kv.y;
        aff_t<F> a;
        load_aff(a, pts, v & 0x7fffffffu);
        // an all-zero affine point is infinity and contributes nothing
        // (Pointxyzz::dadd with an all-zero affine input, point.cpp:2036-2039)
        if (!aff_is_inf(a)) {
            if (v >> 31) fp_neg(a.y, a.y);
            xyzz_madd(acc, a);
        }
    }
    const bool after = end < M && sorted[end].x == cur;
    if (first && before) {
        slot_key[2 * chunk] = cur;
        store_xyzz(&slots[2 * chunk], acc);
    } else if (after) {
        slot_key[2 * chunk + 1] = cur;
        store_xyzz(&slots[2 * chunk + 1], acc);
        const uint32_t m = (off[cur] + cnt[cur] - 1) / T - chunk + 1;   // pieces
        if (m > 2) list_bucket(lists, m, cur);
    } else {
        store_xyzz(&buckets[cur], acc);
    }
}

template <class F>
void launch_k_accumulate(dim3 grid, dim3 block, hipStream_t stream, const uint2* __restrict__ sorted, const uint32_t* __restrict__ total, const job_t* __restrict__ jobs, const uint32_t* __restrict__ kbase, uint32_t njobs, xyzz_t<F>* __restrict__ buckets, uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots, uint32_t T, uint32_t nchunks, const uint32_t* __restrict__ prev, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, merge_lists_t lists) {
    hipLaunchKernelGGL(k_accumulate<F>, grid, block, 0, stream, sorted, total, jobs, kbase, njobs, buckets, slot_key, slots, T, nchunks, prev, off, cnt, lists);
}

template void launch_k_accumulate<fp>(dim3, dim3, hipStream_t, const uint2*, const uint32_t*, const job_t*, const uint32_t*, uint32_t, xyzz*, uint32_t*, xyzz*, uint32_t, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, merge_lists_t);
template void launch_k_accumulate<fp2>(dim3, dim3, hipStream_t, const uint2*, const uint32_t*, const job_t*, const uint32_t*, uint32_t, xyzz2*, uint32_t*, xyzz2*, uint32_t, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, merge_lists_t);

}  // namespace bls
