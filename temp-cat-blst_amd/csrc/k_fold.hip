// Per-job window fold on the device (batched mode).
#include "msm_common.hip.h"
#include "ec_quad.hip.h"
#include "launch.h"

namespace bls {

// one quad per job (four-lane additions and doublings): sum_w 2^(c*w)
// W[gbase + w], Horner from the top window.  Fewer than four jobs still run
// on four quads (16 lanes), the extra ones repeating a job without storing.
template <class F>
__global__ void __launch_bounds__(256)
k_fold(const job_t* __restrict__ jobs, uint32_t njobs, const xyzz_t<F>* __restrict__ W, xyzz_t<F>* __restrict__ out) {
    const uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    if (njobs == 0 || q >= max(njobs, 4u)) return;
    const uint32_t j = q % njobs;
    const job_t& J = jobs[j];
    const int nwin = (int)J.nwin;
    const uint32_t c = J.c;
    F acc;
    fp_zero(acc);
    #pragma unroll 1
    for (int win = nwin - 1; win >= 0; win--) {
        if (win != nwin - 1)
            #pragma unroll 1
            for (uint32_t k = 0; k < c; k++) quad_dbl(acc, acc);
        F w;
        quad_load(w, &W[J.gbase + win]);
        quad_add(acc, acc, w);
    }
    fp_canon(acc, acc);
    if (q < njobs) quad_store(&out[j], acc);
}

template <class F>
void launch_k_fold(hipStream_t stream, const job_t* __restrict__ jobs, uint32_t njobs, const xyzz_t<F>* __restrict__ W, xyzz_t<F>* __restrict__ out) {
    const uint32_t quads = njobs < 4 ? 4 : njobs;
    hipLaunchKernelGGL(k_fold<F>, dim3((quads * 4 + 255) / 256), dim3(256), 0, stream, jobs, njobs, W, out);
}

template void launch_k_fold<fp>(hipStream_t, const job_t*, uint32_t, const xyzz*, xyzz*);
template void launch_k_fold<fp2>(hipStream_t, const job_t*, uint32_t, const xyzz2*, xyzz2*);

}  // namespace bls
