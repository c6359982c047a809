// Bucket reduction sum_b (b+1) S_b per window (point_integrate_buckets,
// blst/scalar/curve/multi_scalar.cpp:730-751).
//
// The B buckets of a window are cut into P = B/L segments of L buckets (L a
// power of two).  Segment s yields the node
//     A_s = sum_t (t+1) S_(sL+t)           (running sum, high to low)
//     V_s = L * sum_t S_(sL+t)             (log2 L doublings)
// and a node for segments [a, a+n) is A = sum_s (U_s + (s-a) L T_s) with
// V = n L sum_s T_s.  Adjacent nodes merge as
//     A = A_left + (A_right + V_right),   V = 2 (V_left + V_right),
// so the root's A is the window sum.  The merges run as a tree of four-lane
// additions: per workgroup 64 nodes in LDS, as many passes as P needs.
// G1 computes in the 28-bit limb field (tail_t<fp> = fq, fq_point.hip.h),
// loading and storing the 32-bit layout; G2 in Fp2.
#include "msm_common.hip.h"
#include "fq_point.hip.h"
#include "launch.h"

namespace bls {

// Segment running sums: one lane per segment.  The running sum (acc += S,
// tot += acc) shares one addition site and the doublings of V one doubling
// site, which keeps the kernel small.
template <class F>
__global__ void __launch_bounds__(256)
k_reduce_segments(const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase,
                  uint32_t ngroups, uint32_t total_segments, const xyzz_t<F>* __restrict__ buckets,
                  const uint32_t* __restrict__ cnt, xyzz_t<F>* __restrict__ nodes) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total_segments) return;
    const uint32_t g = find_owner(sbase, ngroups, t);
    const group_t G = groups[g];
    const uint32_t s = t - G.sb;
    const size_t base = (size_t)G.kb + (size_t)s * G.L;
    xyzz_t<tail_t<F>> acc, tot, x, y;
    xyzz_set_inf(acc);
    xyzz_set_inf(tot);
#pragma unroll 1
    for (uint32_t step = 0; step < 2 * G.L; step++) {
        if (step & 1) {
            x = tot;
            y = acc;
        } else {
            x = acc;
            const size_t key = base + (G.L - 1 - (step >> 1));
            if (cnt[key]) load_xyzz(y, &buckets[key]);   // empty buckets hold no data
            else xyzz_set_inf(y);
        }
        xyzz_add(x, x, y);
        if (step & 1) tot = x;
        else acc = x;
    }
    store_xyzz(&nodes[2 * (size_t)t], tot);
#pragma unroll 1
    for (uint32_t l = G.L; l > 1; l >>= 1) xyzz_dbl_inl(acc, acc);
    store_xyzz(&nodes[2 * (size_t)t + 1], acc);
}

// The same segment sums with one quad per segment (four-lane additions):
// the G2 form, whose one-lane version would hold four Fp2 points per lane.
template <class F>
__global__ void __launch_bounds__(256)
k_reduce_segments_quad(const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase,
                       uint32_t ngroups, uint32_t total_segments, const xyzz_t<F>* __restrict__ buckets,
                       const uint32_t* __restrict__ cnt, xyzz_t<F>* __restrict__ nodes) {
    const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;   // uniform per quad
    if (t >= total_segments) return;
    const uint32_t g = find_owner(sbase, ngroups, t);
    const group_t G = groups[g];
    const uint32_t s = t - G.sb;
    const size_t base = (size_t)G.kb + (size_t)s * G.L;
    F acc, tot, y;
    fp_zero(acc);
    fp_zero(tot);
#pragma unroll 1
    for (uint32_t b = G.L; b-- > 0;) {
        const size_t key = base + b;
        if (cnt[key]) quad_load(y, &buckets[key]);   // empty buckets hold no data
        else fp_zero(y);
        quad_add(acc, acc, y);
        quad_add(tot, tot, acc);
    }
    quad_store(&nodes[2 * (size_t)t], tot);
#pragma unroll 1
    for (uint32_t l = G.L; l > 1; l >>= 1) quad_dbl(acc, acc);
    quad_store(&nodes[2 * (size_t)t + 1], acc);
}

// One workgroup merges n <= 64 consecutive nodes of a group into one.  Quads
// 0-31 (waves 0-1) compute the A of the merges and quads 32-63 (waves 2-3)
// the V, so every wave follows one instruction stream; a level with fewer
// than four merges repeats them on four quads (16 lanes) and stores one.
// tn[g] = (first input node, input nodes, first output node); blk[g] = first
// block of group g.  last: store the root's A to W[g] instead.
// four workgroups (four waves per SIMD) for G1, so that a first pass of
// 1024 workgroups is resident at once
template <class F> constexpr int kTreeBlocksPerCU = 1;
template <> constexpr int kTreeBlocksPerCU<fp> = 4;

template <class F>
__global__ void __launch_bounds__(256, kTreeBlocksPerCU<F>)
k_reduce_tree(const uint32_t* __restrict__ blk, const uint4* __restrict__ tn, uint32_t ngroups,
              const xyzz_t<F>* __restrict__ in, xyzz_t<F>* __restrict__ out, xyzz_t<F>* __restrict__ W, uint32_t last) {
    using E = tail_t<F>;
    __shared__ E sA[kTreeNodes][4], sV[kTreeNodes][4];
    const uint32_t g = find_owner(blk, ngroups, blockIdx.x);
    const uint32_t j = blockIdx.x - blk[g];
    const uint4 t = tn[g];
    const uint32_t first = t.x + kTreeNodes * j;
    const uint32_t n = min(kTreeNodes, t.y - kTreeNodes * j);
    const uint32_t u = threadIdx.x >> 2, ln = threadIdx.x & 3;
    const bool beta = u >= 32;
    const uint32_t v = u & 31;
    if (u < n) {
        E a, b;
        quad_load(a, &in[2 * (size_t)(first + u)]);
        quad_load(b, &in[2 * (size_t)(first + u) + 1]);
        sA[u][ln] = a;
        sV[u][ln] = b;
    }
    __syncthreads();
    for (uint32_t m = n >> 1; m >= 1; m >>= 1) {
        E r;
        if (v < max(m, 4u)) {
            const uint32_t mm = v % m;
            E x, y;
            if (!beta) {
                x = sA[2 * mm + 1][ln];
                y = sV[2 * mm + 1][ln];
                quad_add(x, x, y);
                y = sA[2 * mm][ln];
                quad_add(r, y, x);
            } else {
                x = sV[2 * mm][ln];
                y = sV[2 * mm + 1][ln];
                quad_add(x, x, y);
                quad_dbl(r, x);
            }
        }
        __syncthreads();
        if (v < m) (beta ? sV[v] : sA[v])[ln] = r;
        __syncthreads();
    }
    if (last) {
        if (u == 0) {
            E c;
            fp_canon(c, sA[0][ln]);
            quad_store(&W[g], c);
        }
    } else if (u == 0) {
        quad_store(&out[2 * (size_t)(t.z + j)], sA[0][ln]);
        quad_store(&out[2 * (size_t)(t.z + j) + 1], sV[0][ln]);
    }
}

// G1: one lane per segment (grid and block from the caller); G2: one quad
// per segment
template <class F>
void launch_k_reduce_segments(dim3 grid, dim3 block, hipStream_t stream, const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase, uint32_t ngroups, uint32_t total_segments, const xyzz_t<F>* __restrict__ buckets, const uint32_t* __restrict__ cnt, xyzz_t<F>* __restrict__ nodes) {
    if constexpr (sizeof(F) == sizeof(fp2)) {
        const uint32_t threads = 4 * total_segments;
        hipLaunchKernelGGL(k_reduce_segments_quad<F>, dim3((threads + 255) / 256), dim3(256), 0, stream, groups, sbase, ngroups, total_segments, buckets, cnt, nodes);
    } else {
        hipLaunchKernelGGL(k_reduce_segments<F>, grid, block, 0, stream, groups, sbase, ngroups, total_segments, buckets, cnt, nodes);
    }
}

template <class F>
void launch_k_reduce_tree(dim3 grid, hipStream_t stream, const uint32_t* __restrict__ blk, const uint4* __restrict__ tn, uint32_t ngroups, const xyzz_t<F>* __restrict__ in, xyzz_t<F>* __restrict__ out, xyzz_t<F>* __restrict__ W, uint32_t last) {
    hipLaunchKernelGGL(k_reduce_tree<F>, grid, dim3(256), 0, stream, blk, tn, ngroups, in, out, W, last);
}

#define BLS_REDUCE_INSTANCES(F)                                                                                     \
    template void launch_k_reduce_segments<F>(dim3, dim3, hipStream_t, const group_t*, const uint32_t*, uint32_t,  \
                                              uint32_t, const xyzz_t<F>*, const uint32_t*, xyzz_t<F>*);           \
    template void launch_k_reduce_tree<F>(dim3, hipStream_t, const uint32_t*, const uint4*, uint32_t,              \
                                          const xyzz_t<F>*, xyzz_t<F>*, xyzz_t<F>*, uint32_t);
BLS_REDUCE_INSTANCES(fp)
BLS_REDUCE_INSTANCES(fp2)

}  // namespace bls
