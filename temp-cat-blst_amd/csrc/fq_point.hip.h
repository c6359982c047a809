// The G1 bucket reduction (segment sums and trees, k_reduce.hip) computes in
// the 28-bit limb field of fq.hip.h: a product there is 493 VALU
// instructions against 624, a square 414 against 624.  The partial merges
// stay in the 32-bit form: most of them add two stored pieces once, and the
// conversions below would cost about what the field saves.
//
// The XYZZ formulas of ec.hip.h and ec_quad.hip.h are templates over the
// coordinate type and call fp_mul, fp_sub, ... by overload.  The overloads
// below give them fq with the contract of the 32-bit layer: every input and
// output normalized (limbs below 2^28) and below 2p, zero tested as 0 or p.
// Under that contract every bound fq.hip.h states holds:
//   products     inputs below 2p: output below 4p^2 / 2^392 + p < 1.01p
//   add, dbl     limbs below 2^29, value below 4p: fq_reduce to [0, 2p)
//   sub          a - b + 4p (b normalized below 2p): limbs below 2^29.6,
//                value below 6p: fq_reduce
//   neg          4p - a: fq_reduce
// so the formulas run unchanged; only loads and stores convert.
//
// Stored points keep the 32-bit layout (twelve words per coordinate, values
// below 2p, blst Montgomery form with R = 2^384):
//   load   a word image w (element v = w 2^-384) shifted left by 8 bits is
//          w 2^8 = v 2^392, the exact fq image of v, below 2^390;
//          fq_reduce_wide brings it below 1.01p.
//   store  an fq value c of element v, read as words, is v 2^8.  X and ZZ
//          are stored as they are and Y and ZZZ times 16, so the stored point
//          reads as (2^8 X, 2^12 Y, 2^8 ZZ, 2^12 ZZZ): the same point,
//          rescaled by lambda = 2^4.  Stored coordinates are canonical, so
//          infinity (ZZ = 0) stores ZZ = 0.
// The G2 stages keep their Fp2 arithmetic: tail_t<fp2> is fp2.
#pragma once
#include "ec_quad.hip.h"
#include "fq.hip.h"

namespace bls {

template <class F> struct tail_field { using type = F; };
template <> struct tail_field<fp> { using type = fq; };
template <class F> using tail_t = typename tail_field<F>::type;

__device__ __forceinline__ void fp_zero(fq& r) { fq_zero(r); }
__device__ __forceinline__ bool fp_is_zero(const fq& a) { return fq_is_zero(a); }
__device__ __forceinline__ void fp_canon(fq& r, const fq& a) { fq_canon(r, a); }
__device__ __forceinline__ void fp_mul(fq& r, const fq& a, const fq& b) { fq_mul(r, a, b); }
__device__ __forceinline__ void fp_sqr(fq& r, const fq& a) { fq_sqr(r, a); }
__device__ __forceinline__ void fp_mul_sum(fq& r, const fq& a, const fq& b, const fq& c, const fq& d) {
    fq_mul_sum(r, a, b, c, d);
}

__device__ __forceinline__ void fp_add(fq& r, const fq& a, const fq& b) {
    fq t;
#pragma unroll
    for (int i = 0; i < 14; i++) t.l[i] = a.l[i] + b.l[i];
    fq_reduce(r, t);
}

__device__ __forceinline__ void fp_dbl(fq& r, const fq& a) { fp_add(r, a, a); }

__device__ __forceinline__ void fp_sub(fq& r, const fq& a, const fq& b) {
    fq t;
    fq_sub_lazy(t, a, b);
    fq_reduce(r, t);
}

__device__ __forceinline__ void fp_neg(fq& r, const fq& a) {
    fq t;
    fq_neg_lazy(t, a);
    fq_reduce(r, t);
}

// four-lane moves and selects (ec_quad.hip.h) on fourteen limbs
template <int CTRL>
__device__ __forceinline__ void fp_qmov(fq& r, const fq& a) {
#pragma unroll
    for (int k = 0; k < 14; k++) r.l[k] = qmov<CTRL>(a.l[k]);
}

template <int SRC>
__device__ __forceinline__ void fp_qbcast(fq& r, const fq& a) {
    fp_qmov<quad_perm(SRC, SRC, SRC, SRC)>(r, a);
}

__device__ __forceinline__ void fp_sel(fq& r, bool c, const fq& a, const fq& b) {
#pragma unroll
    for (int k = 0; k < 14; k++) r.l[k] = c ? a.l[k] : b.l[k];
}

// ---------------------------------------------------- loads and stores

// element of the 32-bit word image w (below 2p), exact, below 1.01p
__device__ __forceinline__ void fq_load_words(fq& r, const uint32_t w[12]) {
    fq t;
    fq_from_words<8>(t, w);
    fq_reduce_wide(r, t);
}

__device__ __forceinline__ void fq_load_uint4(fq& r, const uint4* __restrict__ s) {
    uint32_t w[12];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 v = s[k];
        w[4 * k] = v.x; w[4 * k + 1] = v.y; w[4 * k + 2] = v.z; w[4 * k + 3] = v.w;
    }
    fq_load_words(r, w);
}

// canonical words of c, times 16 first for Y and ZZZ (scale16)
__device__ __forceinline__ void fq_store_uint4(uint4* __restrict__ d, const fq& c, bool scale16) {
    fq t, u;
    fq_times4(t, c);
    fq_times4(t, t);
    fp_sel(u, scale16, t, c);
    fq_canon(u, u);
    uint32_t w[12];
    fq_to_words(w, u);
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

__device__ __forceinline__ void load_xyzz(xyzz_t<fq>& p, const xyzz* __restrict__ src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    fq_load_uint4(p.x, s);
    fq_load_uint4(p.y, s + 3);
    fq_load_uint4(p.zz, s + 6);
    fq_load_uint4(p.zzz, s + 9);
}

__device__ __forceinline__ void store_xyzz(xyzz* __restrict__ dst, const xyzz_t<fq>& p) {
    uint4* d = reinterpret_cast<uint4*>(dst);
    fq_store_uint4(d, p.x, false);
    fq_store_uint4(d + 3, p.y, true);
    fq_store_uint4(d + 6, p.zz, false);
    fq_store_uint4(d + 9, p.zzz, true);
}

// this lane's coordinate (X, Y, ZZ or ZZZ by quad lane), as quad_load
__device__ __forceinline__ void quad_load(fq& r, const xyzz* __restrict__ p) {
    fq_load_uint4(r, reinterpret_cast<const uint4*>(p) + 3 * quad_lane());
}

__device__ __forceinline__ void quad_store(xyzz* __restrict__ p, const fq& r) {
    const uint32_t ln = quad_lane();
    fq_store_uint4(reinterpret_cast<uint4*>(p) + 3 * ln, r, (ln & 1) != 0);
}

}  // namespace bls
