// Generator multiples from a comb table and XYZZ -> affine conversion.
#include "msm_common.hip.h"
#include "launch.h"

namespace bls {

// [k_i]G from a comb table T[m][j-1] = [j * 2^(8m)] G (m < nbytes, j < 256):
// one mixed addition per nonzero scalar byte.
__global__ void __launch_bounds__(256)
k_fixed_base(const uint8_t* __restrict__ scalars, uint32_t n, uint32_t nbytes, uint32_t nbits,
             const uint4* __restrict__ table, xyzz* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k[9];
    load_scalar(k, scalars + (size_t)i * nbytes, nbytes, nbits, false);
    xyzz acc;
    xyzz_set_inf(acc);
    #pragma unroll 1
    for (uint32_t m = 0; m < nbytes; m++) {
        const uint32_t b = scalar_bits(k, 8 * m, 8);
        if (b) {
            aff a;
            load_aff(a, table, m * 255 + (b - 1));
            xyzz_madd(acc, a);
        }
    }
    store_xyzz(&out[i], acc);
}

// XYZZ -> affine, Q points per lane sharing one inversion of the product of
// ZZ*ZZZ (= Z^5); prefix products are parked in the output's x field.
template <class F>
__global__ void __launch_bounds__(256)
k_to_affine(const xyzz_t<F>* __restrict__ in, aff_t<F>* __restrict__ out, uint32_t n, uint32_t Q) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t begin = lane * Q;
    if (begin >= n) return;
    const uint32_t end = min(begin + Q, n);
    F acc, t;
    fp_one(acc);
    for (uint32_t i = begin; i < end; i++) {
        xyzz_t<F> p;
        load_xyzz(p, &in[i]);
        out[i].x = acc;
        if (!xyzz_is_inf(p)) {
            fp_mul(t, p.zz, p.zzz);
            fp_mul(acc, acc, t);
        }
    }
    F ia;
    fp_inv(ia, acc);
    for (uint32_t i = end; i-- > begin;) {
        xyzz_t<F> p;
        load_xyzz(p, &in[i]);
        aff_t<F> o;
        if (xyzz_is_inf(p)) {
            fp_zero(o.x);
            fp_zero(o.y);
        } else {
            F it, izz, izzz;
            const F pre = out[i].x;
            fp_mul(t, p.zz, p.zzz);
            fp_mul(it, ia, pre);
            fp_mul(ia, ia, t);
            fp_mul(izz, it, p.zzz);
            fp_mul(izzz, it, p.zz);
            fp_mul(o.x, p.x, izz);
            fp_mul(o.y, p.y, izzz);
            fp_canon(o.x, o.x);
            fp_canon(o.y, o.y);
        }
        out[i] = o;
    }
}

void launch_k_fixed_base(dim3 grid, dim3 block, hipStream_t stream, const uint8_t* __restrict__ scalars, uint32_t n, uint32_t nbytes, uint32_t nbits, const uint4* __restrict__ table, xyzz* __restrict__ out) {
    hipLaunchKernelGGL(k_fixed_base, grid, block, 0, stream, scalars, n, nbytes, nbits, table, out);
}

template <class F>
void launch_k_to_affine(dim3 grid, dim3 block, hipStream_t stream, const xyzz_t<F>* __restrict__ in,
                        aff_t<F>* __restrict__ out, uint32_t n, uint32_t Q) {
    hipLaunchKernelGGL(k_to_affine<F>, grid, block, 0, stream, in, out, n, Q);
}

template void launch_k_to_affine<fp>(dim3, dim3, hipStream_t, const xyzz*, aff*, uint32_t, uint32_t);
template void launch_k_to_affine<fp2>(dim3, dim3, hipStream_t, const xyzz2*, aff2*, uint32_t, uint32_t);

}  // namespace bls
