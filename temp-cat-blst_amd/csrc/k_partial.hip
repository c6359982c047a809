// Merging the partial sums of buckets that span accumulate chunks.
//
// Chunk k leaves the bucket that starts in it and continues in slot 2k+1 and
// the bucket that started before it in slot 2k (partial j = 0 of a bucket
// whose first chunk is k0 is slot 2*k0+1, partial j >= 1 is slot 2*(k0+j)).
// Most spanning buckets cross one boundary (two partials): one lane adds
// them.  Longer spans go to a list and one workgroup per listed bucket sums
// its partials with four-lane additions, first in strides per quad, then as
// a tree in LDS.
#include "msm_common.hip.h"
#include "ec_quad.hip.h"
#include "launch.h"

namespace bls {

__global__ void __launch_bounds__(256)
k_partial_light(const uint32_t* __restrict__ slot_key, const xyzz* __restrict__ slots,
                const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nchunks,
                xyzz* __restrict__ buckets, uint32_t* __restrict__ heavy_count, uint32_t* __restrict__ heavy) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nchunks) return;
    const uint32_t key = slot_key[2 * k + 1];
    if (key == kInvalid) return;
    const uint32_t k1 = (off[key] + cnt[key] - 1) / T;
    if (k1 == k + 1) {
        xyzz a, b;
        load_xyzz(a, &slots[2 * k + 1]);
        load_xyzz(b, &slots[2 * k + 2]);
        xyzz_add(a, a, b);
        store_xyzz(&buckets[key], a);
    } else {
        heavy[atomicAdd(heavy_count, 1u)] = key;
    }
}

constexpr uint32_t kHeavyThreads = 1024;
constexpr uint32_t kHeavyQuads = kHeavyThreads / 4;

// One listed bucket per workgroup iteration; the grid strides over the list.
__global__ void __launch_bounds__(kHeavyThreads)
k_partial_heavy(const uint32_t* __restrict__ heavy_count, const uint32_t* __restrict__ heavy,
                const xyzz* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                uint32_t T, xyzz* __restrict__ buckets) {
    __shared__ xyzz node[kHeavyQuads];
    const uint32_t u = threadIdx.x >> 2, ln = threadIdx.x & 3;
    const uint32_t nheavy = *heavy_count;
    for (uint32_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const uint32_t key = heavy[h];
        const uint32_t k0 = off[key] / T;
        const uint32_t m = (off[key] + cnt[key] - 1) / T - k0 + 1;   // partials
        fp acc;
        fp_zero(acc);
        for (uint32_t j = u; j < m; j += kHeavyQuads) {
            fp x;
            quad_load(x, &slots[j == 0 ? 2 * k0 + 1 : 2 * (k0 + j)]);
            quad_add(acc, acc, x);
        }
        reinterpret_cast<fp*>(&node[u])[ln] = acc;
        __syncthreads();
        // pairwise tree; at least four quads (16 lanes) run each level, the
        // extra ones repeating a merge without storing it
        for (uint32_t half = kHeavyQuads / 2; half >= 1; half >>= 1) {
            fp r;
            const bool run = u < max(half, 4u);
            if (run) {
                const uint32_t v = u % half;
                fp a, b;
                a = reinterpret_cast<const fp*>(&node[v])[ln];
                b = reinterpret_cast<const fp*>(&node[v + half])[ln];
                quad_add(r, a, b);
            }
            __syncthreads();
            if (u < half) reinterpret_cast<fp*>(&node[u])[ln] = r;
            __syncthreads();
        }
        if (u == 0) quad_store(&buckets[key], reinterpret_cast<const fp*>(&node[0])[ln]);
        __syncthreads();
    }
}

void launch_k_partial_light(dim3 grid, dim3 block, hipStream_t stream, const uint32_t* __restrict__ slot_key, const xyzz* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nchunks, xyzz* __restrict__ buckets, uint32_t* __restrict__ heavy_count, uint32_t* __restrict__ heavy) {
    hipLaunchKernelGGL(k_partial_light, grid, block, 0, stream, slot_key, slots, off, cnt, T, nchunks, buckets, heavy_count, heavy);
}

void launch_k_partial_heavy(dim3 grid, hipStream_t stream, const uint32_t* __restrict__ heavy_count, const uint32_t* __restrict__ heavy, const xyzz* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, xyzz* __restrict__ buckets) {
    hipLaunchKernelGGL(k_partial_heavy, grid, dim3(kHeavyThreads), 0, stream, heavy_count, heavy, slots, off, cnt, T, buckets);
}

}  // namespace bls
