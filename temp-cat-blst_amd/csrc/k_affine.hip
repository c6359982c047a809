// Jacobian -> affine batch normalisation of G1 and G2 points
// (points_to_affine, blst/scalar/curve/multi_scalar.cpp:160-238): x = X / Z^2,
// y = Y / Z^3 in canonical Montgomery form, and the all-zero affine encoding
// for Z == 0.
//
// Montgomery's trick per lane: lane k of T handles points k, k + T, k + 2T,
// ... (neighbouring lanes read neighbouring points, so a wave's loads cover
// one contiguous stretch), multiplies their Z into a running product whose
// prefixes wait in the output's x field, inverts the product once, and walks
// back.  Per point: one multiplication forward and six back; per lane: one
// inversion.
#include "msm_common.hip.h"
#include "launch.h"

namespace bls {

__device__ __forceinline__ void load_fp_u4(fp& r, const uint4* __restrict__ s) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 v = s[k];
        r.l[4 * k] = v.x; r.l[4 * k + 1] = v.y; r.l[4 * k + 2] = v.z; r.l[4 * k + 3] = v.w;
    }
}

__device__ __forceinline__ void load_fp_u4(fp2& r, const uint4* __restrict__ s) {
    load_fp_u4(r.c0, s);
    load_fp_u4(r.c1, s + 3);
}

__device__ __forceinline__ void store_fp_u4(uint4* __restrict__ d, const fp& a) {
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = make_uint4(a.l[4 * k], a.l[4 * k + 1], a.l[4 * k + 2], a.l[4 * k + 3]);
}

__device__ __forceinline__ void store_fp_u4(uint4* __restrict__ d, const fp2& a) {
    store_fp_u4(d, a.c0);
    store_fp_u4(d + 3, a.c1);
}

// coordinate c (0 = X, 1 = Y, 2 = Z) of Jacobian point i (blst_p1: 144
// bytes, blst_p2: 288 bytes)
template <class F>
__device__ __forceinline__ void load_jac_coord(F& r, const uint4* __restrict__ pts, uint32_t i, int c) {
    constexpr int W = sizeof(F) / 16;   // uint4 per coordinate
    load_fp_u4(r, pts + (size_t)i * 3 * W + W * c);
}

// in: n Jacobian points (X, Y, Z), coordinates below p; out: n affine points.
// T lanes (T <= n); in and out must not overlap.
template <class F>
__global__ void __launch_bounds__(256)
k_jac_to_affine(const uint4* __restrict__ in, aff_t<F>* __restrict__ out, uint32_t n, uint32_t T) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= T) return;
    F acc;
    fp_one(acc);
    uint32_t last = k;
    for (uint32_t i = k; i < n; i += T) {
        F z;
        load_jac_coord(z, in, i, 2);
        store_fp_u4(reinterpret_cast<uint4*>(&out[i].x), acc);
        if (!fp_is_zero(z)) fp_mul(acc, acc, z);
        last = i;
    }
    F ia;
    fp_inv(ia, acc);   // ia = 1 / (product of this lane's nonzero Z)
    for (uint32_t i = last;; i -= T) {
        F z, pre, x, y;
        load_jac_coord(z, in, i, 2);
        load_fp_u4(pre, reinterpret_cast<const uint4*>(&out[i].x));
        if (fp_is_zero(z)) {
            fp_zero(x);
            fp_zero(y);
        } else {
            F iz, iz2, iz3;
            fp_mul(iz, ia, pre);    // 1 / Z_i
            fp_mul(ia, ia, z);      // 1 / (product of the Z before i)
            fp_mul(iz2, iz, iz);
            fp_mul(iz3, iz2, iz);
            load_jac_coord(x, in, i, 0);
            load_jac_coord(y, in, i, 1);
            fp_mul(x, x, iz2);
            fp_mul(y, y, iz3);
            fp_canon(x, x);
            fp_canon(y, y);
        }
        store_fp_u4(reinterpret_cast<uint4*>(&out[i].x), x);
        store_fp_u4(reinterpret_cast<uint4*>(&out[i].y), y);
        if (i < T) break;
    }
}

template <class F>
void launch_k_jac_to_affine(hipStream_t stream, const uint4* __restrict__ in, aff_t<F>* __restrict__ out, uint32_t n,
                            uint32_t T) {
    hipLaunchKernelGGL(k_jac_to_affine<F>, dim3((T + 255) / 256), dim3(256), 0, stream, in, out, n, T);
}

template void launch_k_jac_to_affine<fp>(hipStream_t, const uint4*, aff*, uint32_t, uint32_t);
template void launch_k_jac_to_affine<fp2>(hipStream_t, const uint4*, aff2*, uint32_t, uint32_t);

}  // namespace bls
