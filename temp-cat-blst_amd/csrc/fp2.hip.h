// Fp2 = Fp[i] / (i^2 + 1) on gfx950, the field of G2 (blst_fp2,
// bindings/blst_types.h:100; blst/scalar/fields/Fp2.h).  An element is
// (c0, c1), each an Fp in the [0, 2p) form of fp.hip.h; the functions are
// overloads of the Fp ones, so the point formulas are written once for both
// fields.
#pragma once
#include "fp.hip.h"

namespace bls {

struct fp2 {
    fp c0, c1;
};

__device__ __forceinline__ void fp_zero(fp2& r) { fp_zero(r.c0); fp_zero(r.c1); }
__device__ __forceinline__ void fp_one(fp2& r) { fp_one(r.c0); fp_zero(r.c1); }
__device__ __forceinline__ bool fp_is_zero(const fp2& a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
__device__ __forceinline__ void fp_canon(fp2& r, const fp2& a) { fp_canon(r.c0, a.c0); fp_canon(r.c1, a.c1); }
__device__ __forceinline__ void fp_add(fp2& r, const fp2& a, const fp2& b) {
    fp_add(r.c0, a.c0, b.c0);
    fp_add(r.c1, a.c1, b.c1);
}
__device__ __forceinline__ void fp_sub(fp2& r, const fp2& a, const fp2& b) {
    fp_sub(r.c0, a.c0, b.c0);
    fp_sub(r.c1, a.c1, b.c1);
}
__device__ __forceinline__ void fp_dbl(fp2& r, const fp2& a) { fp_add(r, a, a); }
__device__ __forceinline__ void fp_neg(fp2& r, const fp2& a) { fp_neg(r.c0, a.c0); fp_neg(r.c1, a.c1); }

// c0 = a0 b0 - a1 b1 = a0 b0 + a1 (2p - b1), c1 = a0 b1 + a1 b0: two fused
// sums of products, each with one Montgomery reduction (fp_mul_sum), about
// 1800 VALU instructions against ~2100 for Karatsuba's three products and
// five additions; inputs below 2p give results below 2p
__device__ __forceinline__ void fp_mul(fp2& r, const fp2& a, const fp2& b) {
    fp nb1, c0;
    fp_neg(nb1, b.c1);
    fp_mul_sum(c0, a.c0, b.c0, a.c1, nb1);
    fp_mul_sum(r.c1, a.c0, b.c1, a.c1, b.c0);
    r.c0 = c0;
}

// a*b + c*d as two Fp2 products and a sum
__device__ __forceinline__ void fp_mul_sum(fp2& r, const fp2& a, const fp2& b, const fp2& c, const fp2& d) {
    fp2 u, v;
    fp_mul(u, a, b);
    fp_mul(v, c, d);
    fp_add(r, u, v);
}

// two Fp products: c0 = (a0 + a1)(a0 - a1), c1 = 2 a0 a1
__device__ __forceinline__ void fp_sqr(fp2& r, const fp2& a) {
    fp s, d, m;
    fp_add(s, a.c0, a.c1);
    fp_sub(d, a.c0, a.c1);
    fp_mul(m, a.c0, a.c1);
    fp_mul(r.c0, s, d);
    fp_dbl(r.c1, m);
}

// 1 / (a0 + a1 i) = (a0 - a1 i) / (a0^2 + a1^2), one Fp inversion; 0 for 0
// (a0^2 + a1^2 by fp_mul_sum is below 1.9p, which fp_inv takes)
__device__ __forceinline__ void fp_inv(fp2& r, const fp2& a) {
    fp n, ni, t;
    fp_mul_sum(n, a.c0, a.c0, a.c1, a.c1);
    fp_inv(ni, n);
    fp_mul(r.c0, a.c0, ni);
    fp_neg(t, a.c1);
    fp_mul(r.c1, t, ni);
}

}  // namespace bls
