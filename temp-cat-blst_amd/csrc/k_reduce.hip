// Bucket reduction sum_b b*S_b per window and the window sums.
#include "msm_common.hip.h"
#include "launch.h"

namespace bls {

// Segment s of L buckets of group g (bucket b holds digit b+1):
// H = sum_t (t+1) S_(sL+t) + (sL) * sum_t S_(sL+t), the segment's share of
// sum_b (b+1) S_b.  The running sum (acc += S, tot += acc, high to low), the
// double-and-add of sL * acc and the final tot += r share one addition site
// and one doubling site, which keeps the kernel small.
__global__ void __launch_bounds__(256)
k_reduce_segments(const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase,
                  uint32_t ngroups, uint32_t total_segments, const xyzz* __restrict__ buckets,
                  xyzz* __restrict__ H) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total_segments) return;
    const uint32_t g = find_owner(sbase, ngroups, t);
    const group_t G = groups[g];
    const uint32_t s = t - G.sb;
    const size_t base = (size_t)G.kb + (size_t)s * G.L;
    const uint32_t m = s * G.L;
    const int mbits = m ? 32 - __builtin_clz(m) : 0;
    xyzz acc, tot, r, x, y;
    xyzz_set_inf(acc);
    xyzz_set_inf(tot);
    xyzz_set_inf(r);
    // steps 0 .. 2L-1: running sum; then per bit of m (high to low): double r
    // and, for a set bit, r += acc; last: tot += r
    const int nrun = 2 * (int)G.L;
    int bit = mbits - 1;
    bool dbl_pending = true;
#pragma unroll 1
    for (int step = 0;; step++) {
        int kind;                       // 0 acc+=S, 1 tot+=acc, 2 r+=acc, 3 tot+=r
        if (step < nrun) {
            kind = step & 1;
        } else if (bit >= 0) {
            if (dbl_pending) {
                xyzz_dbl_inl(r, r);
                dbl_pending = false;
                if (!((m >> bit) & 1u)) { bit--; dbl_pending = true; continue; }
            }
            kind = 2;
        } else {
            kind = 3;
        }
        if (kind == 0) {
            load_xyzz(y, &buckets[base + (G.L - 1 - (uint32_t)(step >> 1))]);
            x = acc;
        } else if (kind == 1) {
            x = tot;
            y = acc;
        } else if (kind == 2) {
            x = r;
            y = acc;
        } else {
            x = tot;
            y = r;
        }
        xyzz_add(x, x, y);
        if (kind == 0) acc = x;
        else if (kind == 2) { r = x; bit--; dbl_pending = true; }
        else tot = x;
        if (kind == 3) break;
    }
    store_xyzz(&H[t], tot);
}

// within each group: H[sb + s] += H[sb + s + step] for s a multiple of 2*step
__global__ void __launch_bounds__(256)
k_pair_sum(const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase, uint32_t ngroups,
           uint32_t total_segments, xyzz* __restrict__ H, uint32_t step) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total_segments) return;
    const uint32_t g = find_owner(sbase, ngroups, t);
    const uint32_t s = t - groups[g].sb;
    if ((s % (2 * step)) || s + step >= groups[g].P) return;
    xyzz x, y;
    load_xyzz(x, &H[t]);
    load_xyzz(y, &H[t + step]);
    xyzz_add(x, x, y);
    store_xyzz(&H[t], x);
}

// window sums of every group -> W[g]
__global__ void __launch_bounds__(256)
k_gather_windows(const group_t* __restrict__ groups, uint32_t ngroups, const xyzz* __restrict__ H,
                 xyzz* __restrict__ W) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    xyzz x;
    load_xyzz(x, &H[groups[g].sb]);
    store_xyzz(&W[g], x);
}

void launch_k_reduce_segments(dim3 grid, dim3 block, hipStream_t stream, const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase, uint32_t ngroups, uint32_t total_segments, const xyzz* __restrict__ buckets, xyzz* __restrict__ H) {
    hipLaunchKernelGGL(k_reduce_segments, grid, block, 0, stream, groups, sbase, ngroups, total_segments, buckets, H);
}

void launch_k_pair_sum(dim3 grid, dim3 block, hipStream_t stream, const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase, uint32_t ngroups, uint32_t total_segments, xyzz* __restrict__ H, uint32_t step) {
    hipLaunchKernelGGL(k_pair_sum, grid, block, 0, stream, groups, sbase, ngroups, total_segments, H, step);
}

void launch_k_gather_windows(dim3 grid, dim3 block, hipStream_t stream, const group_t* __restrict__ groups, uint32_t ngroups, const xyzz* __restrict__ H, xyzz* __restrict__ W) {
    hipLaunchKernelGGL(k_gather_windows, grid, block, 0, stream, groups, ngroups, H, W);
}

}  // namespace bls
