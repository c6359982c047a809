// Per-job window fold on the device (batched mode).
#include "msm_common.hip.h"
#include "launch.h"

namespace bls {

// one lane per job: sum_w 2^(c*w) W[gbase + w], Horner from the top window
__global__ void __launch_bounds__(64)
k_fold(const job_t* __restrict__ jobs, uint32_t njobs, const xyzz* __restrict__ W, xyzz* __restrict__ out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const job_t J = jobs[j];
    xyzz acc, w;
    xyzz_set_inf(acc);
    #pragma unroll 1
    for (int win = (int)J.nwin - 1; win >= 0; win--) {
        if (win != (int)J.nwin - 1)
            #pragma unroll 1
            for (uint32_t k = 0; k < J.c; k++) xyzz_dbl_inl(acc, acc);
        load_xyzz(w, &W[J.gbase + win]);
        xyzz_add(acc, acc, w);
    }
    xyzz_canon(acc, acc);
    store_xyzz(&out[j], acc);
}

void launch_k_fold(dim3 grid, dim3 block, hipStream_t stream, const job_t* __restrict__ jobs, uint32_t njobs, const xyzz* __restrict__ W, xyzz* __restrict__ out) {
    hipLaunchKernelGGL(k_fold, grid, block, 0, stream, jobs, njobs, W, out);
}

}  // namespace bls
