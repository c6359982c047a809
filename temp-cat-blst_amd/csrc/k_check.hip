// Self-check of the four-lane point arithmetic: on valid points, every case
// of an addition (general, equal points in different frames, opposite
// points, infinity on either side) and of a doubling must give the same
// coordinates (as field elements) as the one-lane formulas of ec.hip.h.
// The same formulas over the 28-bit limb field (fq_point.hip.h), one-lane
// and four-lane, on the same points loaded from their 32-bit word images,
// must give the same elements again.
#include "msm_common.hip.h"
#include "fq_point.hip.h"
#include "launch.h"

namespace bls {

__device__ __forceinline__ void xyzz_coord(fp& r, const xyzz& p, uint32_t ln) {
    fp_sel(r, ln == 0, p.x, p.y);
    fp_sel(r, ln == 2, p.zz, r);
    fp_sel(r, ln == 3, p.zzz, r);
}

// the two forms may return different representatives in [0, 2p) of one
// field element (the one-lane Y3 shares a reduction between two products),
// so elements are compared in canonical form
__device__ __forceinline__ uint32_t fp_differs(const fp& a, const fp& b) {
    fp ca, cb;
    fp_canon(ca, a);
    fp_canon(cb, b);
    uint32_t d = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) d |= ca.l[k] ^ cb.l[k];
    return d != 0;
}

// an fq value against the element of a 32-bit value
__device__ __forceinline__ uint32_t fq_differs(const fq& a, const fp& b) {
    fq ca, cb;
    fq_load_words(cb, b.l);
    fq_canon(cb, cb);
    fq_canon(ca, a);
    uint32_t d = 0;
#pragma unroll
    for (int k = 0; k < 14; k++) d |= ca.l[k] ^ cb.l[k];
    return d != 0;
}

__device__ __forceinline__ void xyzz_to_fq(xyzz_t<fq>& r, const xyzz& p) {
    fq_load_words(r.x, p.x.l);
    fq_load_words(r.y, p.y.l);
    fq_load_words(r.zz, p.zz.l);
    fq_load_words(r.zzz, p.zzz.l);
}

__device__ __forceinline__ void xyzz_coord(fq& r, const xyzz_t<fq>& p, uint32_t ln) {
    fp_sel(r, ln == 0, p.x, p.y);
    fp_sel(r, ln == 2, p.zz, r);
    fp_sel(r, ln == 3, p.zzz, r);
}

// one quad per point index i: P = pts[i], Q = pts[i + 1 mod n]
__global__ void __launch_bounds__(256)
k_quad_check(const uint4* __restrict__ pts, uint32_t n, uint32_t* __restrict__ bad) {
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    if (i >= n) return;
    const uint32_t ln = threadIdx.x & 3;
    aff P, Q;
    load_aff(P, pts, i);
    load_aff(Q, pts, (i + 1) % n);
    xyzz a, b, c, cn, z, one;
    xyzz_from_aff(a, P);
    xyzz_dbl_inl(a, a);                 // 2P
    xyzz_from_aff(b, Q);
    xyzz_dbl_inl(b, b);                 // 2Q
    xyzz_from_aff(c, P);
    xyzz_from_aff(one, P);
    xyzz_add(c, c, one);                // 2P in the addition's frame
    xyzz_neg(cn, c);                    // -2P
    xyzz_set_inf(z);
    uint32_t miss = 0;
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        xyzz x, y, s;
        switch (k) {
        case 0: x = a; y = b; break;     // general
        case 1: x = a; y = c; break;     // equal points, different frames
        case 2: x = a; y = cn; break;    // opposite points
        case 3: x = z; y = b; break;
        case 4: x = a; y = z; break;
        case 5: x = z; y = z; break;
        case 6: x = a; y = a; break;     // equal coordinates
        default: x = one; y = b; break;  // ZZ = 1 on one side
        }
        xyzz_add(s, x, y);
        fp xl, yl, r, sl;
        xyzz_coord(xl, x, ln);
        xyzz_coord(yl, y, ln);
        quad_add(r, xl, yl);
        xyzz_coord(sl, s, ln);
        miss += fp_differs(r, sl);
        xyzz_t<fq> xq, yq, sq;
        xyzz_to_fq(xq, x);
        xyzz_to_fq(yq, y);
        xyzz_add(sq, xq, yq);
        fq xlq, ylq, rq, slq;
        xyzz_coord(slq, sq, ln);
        miss += fq_differs(slq, sl);
        xyzz_coord(xlq, xq, ln);
        xyzz_coord(ylq, yq, ln);
        quad_add(rq, xlq, ylq);
        miss += fq_differs(rq, sl);
    }
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        xyzz x = k == 0 ? a : (k == 1 ? c : z), s;
        xyzz_dbl_inl(s, x);
        fp xl, r, sl;
        xyzz_coord(xl, x, ln);
        quad_dbl(r, xl);
        xyzz_coord(sl, s, ln);
        miss += fp_differs(r, sl);
        xyzz_t<fq> xq, sq;
        xyzz_to_fq(xq, x);
        xyzz_dbl_inl(sq, xq);
        fq xlq, rq, slq;
        xyzz_coord(slq, sq, ln);
        miss += fq_differs(slq, sl);
        xyzz_coord(xlq, xq, ln);
        quad_dbl(rq, xlq);
        miss += fq_differs(rq, sl);
    }
    if (miss) atomicAdd(bad, miss);
}

void launch_k_quad_check(dim3 grid, dim3 block, hipStream_t stream, const uint4* __restrict__ pts, uint32_t n, uint32_t* __restrict__ bad) {
    hipLaunchKernelGGL(k_quad_check, grid, block, 0, stream, pts, n, bad);
}

}  // namespace bls
