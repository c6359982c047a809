// Jacobian -> affine batch normalisation of G1 points (points_to_affine,
// blst/scalar/curve/multi_scalar.cpp:160-238): x = X / Z^2, y = Y / Z^3 in
// canonical Montgomery form, and the all-zero affine encoding for Z == 0.
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

// coordinate c (0 = X, 1 = Y, 2 = Z) of Jacobian point i (144 bytes each)
__device__ __forceinline__ void load_jac_coord(fp& r, const uint4* __restrict__ pts, uint32_t i, int c) {
    const uint4* s = pts + (size_t)i * 9 + 3 * c;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 v = s[k];
        r.l[4 * k] = v.x; r.l[4 * k + 1] = v.y; r.l[4 * k + 2] = v.z; r.l[4 * k + 3] = v.w;
    }
}

__device__ __forceinline__ void store_fp(fp* __restrict__ dst, const fp& a) {
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = make_uint4(a.l[4 * k], a.l[4 * k + 1], a.l[4 * k + 2], a.l[4 * k + 3]);
}

__device__ __forceinline__ void load_fp(fp& r, const fp* __restrict__ src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 v = s[k];
        r.l[4 * k] = v.x; r.l[4 * k + 1] = v.y; r.l[4 * k + 2] = v.z; r.l[4 * k + 3] = v.w;
    }
}

// in: n Jacobian points (X, Y, Z), coordinates below p; out: n affine points.
// T lanes (T <= n); in and out must not overlap.
__global__ void __launch_bounds__(256)
k_jac_to_affine(const uint4* __restrict__ in, aff* __restrict__ out, uint32_t n, uint32_t T) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= T) return;
    fp acc;
    fp_one(acc);
    uint32_t last = k;
    for (uint32_t i = k; i < n; i += T) {
        fp z;
        load_jac_coord(z, in, i, 2);
        store_fp(&out[i].x, acc);
        if (!fp_is_zero(z)) fp_mul(acc, acc, z);
        last = i;
    }
    fp ia;
    fp_inv(ia, acc);   // ia = 1 / (product of this lane's nonzero Z)
    for (uint32_t i = last;; i -= T) {
        fp z, pre, x, y;
        load_jac_coord(z, in, i, 2);
        load_fp(pre, &out[i].x);
        if (fp_is_zero(z)) {
            fp_zero(x);
            fp_zero(y);
        } else {
            fp iz, iz2, iz3;
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
        store_fp(&out[i].x, x);
        store_fp(&out[i].y, y);
        if (i < T) break;
    }
}

void launch_k_jac_to_affine(hipStream_t stream, const uint4* __restrict__ in, aff* __restrict__ out, uint32_t n,
                            uint32_t T) {
    hipLaunchKernelGGL(k_jac_to_affine, dim3((T + 255) / 256), dim3(256), 0, stream, in, out, n, T);
}

}  // namespace bls
