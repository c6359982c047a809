// Pairwise tree over the partial sums of buckets that span chunks.
#include "msm_common.hip.h"
#include "launch.h"

namespace bls {

// Round with distance d of the pairwise tree over the partials of one bucket:
// partial j (j = 0 in slot 2*k0+1, j >= 1 in slot 2*(k0+j)) absorbs partial
// j+d when j is a multiple of 2d.
__global__ void __launch_bounds__(256)
k_partial_round(const uint32_t* __restrict__ slot_key, xyzz* __restrict__ slots,
                const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                uint32_t T, uint32_t nslots, uint32_t d) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint32_t key = slot_key[s];
    if (key == kInvalid) return;
    const uint32_t k0 = off[key] / T;
    const uint32_t k1 = (off[key] + cnt[key] - 1) / T;
    const uint32_t j = (s & 1) ? 0u : (s >> 1) - k0;
    if (j % (2 * d)) return;
    const uint32_t jj = j + d;
    if (jj > k1 - k0) return;
    xyzz a, b;
    load_xyzz(a, &slots[s]);
    load_xyzz(b, &slots[2 * (k0 + jj)]);
    xyzz_add(a, a, b);
    store_xyzz(&slots[s], a);
}

__global__ void __launch_bounds__(256)
k_partial_final(const uint32_t* __restrict__ slot_key, const xyzz* __restrict__ slots,
                xyzz* __restrict__ buckets, uint32_t nslots) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots || !(s & 1)) return;
    const uint32_t key = slot_key[s];
    if (key == kInvalid) return;
    xyzz a;
    load_xyzz(a, &slots[s]);
    store_xyzz(&buckets[key], a);
}

void launch_k_partial_round(dim3 grid, dim3 block, hipStream_t stream, const uint32_t* __restrict__ slot_key, xyzz* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nslots, uint32_t d) {
    hipLaunchKernelGGL(k_partial_round, grid, block, 0, stream, slot_key, slots, off, cnt, T, nslots, d);
}

void launch_k_partial_final(dim3 grid, dim3 block, hipStream_t stream, const uint32_t* __restrict__ slot_key, const xyzz* __restrict__ slots, xyzz* __restrict__ buckets, uint32_t nslots) {
    hipLaunchKernelGGL(k_partial_final, grid, block, 0, stream, slot_key, slots, buckets, nslots);
}

}  // namespace bls
