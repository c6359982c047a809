// G1 point arithmetic on gfx950 for y^2 = x^3 + 4 over Fp.
//
// Formats (all field elements Montgomery, twelve 32-bit limbs):
//   aff   (x, y); the all-zero pair is the point at infinity, as in
//         blst_p1_affine (bindings/blst_types.h:95-152).
//   xyzz  (X, Y, ZZ, ZZZ) with x = X/ZZ, y = Y/ZZZ; infinity is ZZ = ZZZ = 0.
//         This is the bucket representation of the reference
//         (blst/scalar/curve/point.h:347-384).
//   jac   (X, Y, Z) Jacobian, the blst_p1 layout; infinity is Z = 0.
//
// The same formats over Fp2 are the G2 points (blst_p2_affine, blst_p2); the
// formulas are templates over the field and serve both groups (a = 0 on
// both curves, and b does not appear in additions or doublings).
//
// Formula sources are the EFD entries the reference names
// (blst/scalar/curve/point.cpp:1933-1936, 2021-2025): madd-2008-s and
// mdbl-2008-s-1 for XYZZ+affine, add-2008-s and dbl-2008-s-1 for XYZZ+XYZZ.
// The exceptional inputs (equal points, opposite points, infinity) are
// resolved inside the adders without a second multiplication sequence.
#pragma once
#include "fp.hip.h"
#include "fp2.hip.h"

namespace bls {

template <class F> struct aff_t { F x, y; };
template <class F> struct xyzz_t { F x, y, zz, zzz; };
template <class F> struct jac_t { F x, y, z; };
using aff = aff_t<fp>;
using xyzz = xyzz_t<fp>;
using jac = jac_t<fp>;
using aff2 = aff_t<fp2>;
using xyzz2 = xyzz_t<fp2>;
using jac2 = jac_t<fp2>;

template <class F>
__device__ __forceinline__ void xyzz_set_inf(xyzz_t<F>& p) {
    fp_zero(p.x); fp_zero(p.y); fp_zero(p.zz); fp_zero(p.zzz);
}

template <class F>
__device__ __forceinline__ bool xyzz_is_inf(const xyzz_t<F>& p) { return fp_is_zero(p.zz); }

template <class F>
__device__ __forceinline__ bool aff_is_inf(const aff_t<F>& a) {
    return fp_is_zero(a.x) && fp_is_zero(a.y);
}

template <class F>
__device__ __forceinline__ void xyzz_from_aff(xyzz_t<F>& p, const aff_t<F>& a) {
    p.x = a.x; p.y = a.y; fp_one(p.zz); fp_one(p.zzz);
}

// dbl-2008-s-1: 2*P in XYZZ
template <class F>
__device__ __forceinline__ void xyzz_dbl_inl(xyzz_t<F>& r, const xyzz_t<F>& p) {
    if (xyzz_is_inf(p)) { r = p; return; }
    F U, V, W, S, M, t;
    fp_dbl(U, p.y);
    fp_sqr(V, U);
    fp_mul(W, U, V);
    fp_mul(S, p.x, V);
    fp_sqr(M, p.x);
    fp_dbl(t, M);
    fp_add(M, M, t);         // M = 3X^2 (a = 0)
    xyzz_t<F> o;
    fp_sqr(o.x, M);
    fp_dbl(t, S);
    fp_sub(o.x, o.x, t);
    fp_sub(t, S, o.x);
    fp_mul(t, t, M);
    fp_mul(o.y, W, p.y);
    fp_sub(o.y, t, o.y);
    fp_mul(o.zz, V, p.zz);
    fp_mul(o.zzz, W, p.zzz);
    r = o;
}


// Equal inputs reuse the addition's own multiplication sequence: with
// P = R = 0 (the same point), substituting P' = 2*S1 and R' = 3*U1^2 turns
// add-2008-s into dbl-2008-s-1 of the first input in the addition's scaled
// frame (U1, S1, ZZ1*ZZ2, ZZZ1*ZZZ2), except that X3 = R'^2 - 2Q' carries no
// -PPP' term, which is added back.  P = 0 with R != 0 is P + (-P) = infinity.

// acc += a (a affine, not infinity): madd-2008-s, 8M + 2S; doubling when a == acc
template <class F>
__device__ __forceinline__ void xyzz_madd(xyzz_t<F>& acc, const aff_t<F>& a) {
    if (xyzz_is_inf(acc)) { xyzz_from_aff(acc, a); return; }
    F P, R;
    fp_mul(P, a.x, acc.zz);
    fp_mul(R, a.y, acc.zzz);
    fp_sub(P, P, acc.x);     // P = U2 - X1
    fp_sub(R, R, acc.y);     // R = S2 - Y1
    const bool pz = fp_is_zero(P);
    const bool dbl = pz && fp_is_zero(R);
    if (pz && !dbl) { xyzz_set_inf(acc); return; }
    if (dbl) {
        F t;
        fp_dbl(P, acc.y);        // P' = 2*Y1
        fp_sqr(t, acc.x);
        fp_dbl(R, t);
        fp_add(R, R, t);         // R' = 3*X1^2
    }
    F PP, PPP, Q, t;
    fp_sqr(PP, P);
    fp_mul(PPP, PP, P);
    fp_mul(Q, acc.x, PP);
    fp_sqr(t, R);
    if (!dbl) fp_sub(t, t, PPP);
    fp_sub(t, t, Q);
    fp_sub(acc.x, t, Q);     // X3 = R^2 - PPP - 2Q
    fp_sub(Q, Q, acc.x);
    fp_neg(t, acc.y);
    fp_mul_sum(acc.y, Q, R, t, PPP);   // Y3 = R(Q - X3) - Y1*PPP, one reduction
    fp_mul(acc.zz, acc.zz, PP);
    fp_mul(acc.zzz, acc.zzz, PPP);
}

// r = p + q: add-2008-s, 12M + 2S; doubling when p == q
template <class F>
__device__ __forceinline__ void xyzz_add(xyzz_t<F>& r, const xyzz_t<F>& p, const xyzz_t<F>& q) {
    if (xyzz_is_inf(q)) { r = p; return; }
    if (xyzz_is_inf(p)) { r = q; return; }
    F U1, S1, P, R;
    fp_mul(U1, p.x, q.zz);
    fp_mul(S1, p.y, q.zzz);
    fp_mul(P, q.x, p.zz);
    fp_mul(R, q.y, p.zzz);
    fp_sub(P, P, U1);
    fp_sub(R, R, S1);
    const bool pz = fp_is_zero(P);
    const bool dbl = pz && fp_is_zero(R);
    if (pz && !dbl) { xyzz_set_inf(r); return; }
    if (dbl) {
        F t;
        fp_dbl(P, S1);           // P' = 2*S1
        fp_sqr(t, U1);
        fp_dbl(R, t);
        fp_add(R, R, t);         // R' = 3*U1^2
    }
    F PP, PPP, Q, t;
    xyzz_t<F> o;
    fp_sqr(PP, P);
    fp_mul(PPP, PP, P);
    fp_mul(Q, U1, PP);
    fp_sqr(t, R);
    if (!dbl) fp_sub(t, t, PPP);
    fp_sub(t, t, Q);
    fp_sub(o.x, t, Q);
    fp_sub(Q, Q, o.x);
    fp_neg(t, S1);
    fp_mul_sum(o.y, Q, R, t, PPP);     // Y3 = R(Q - X3) - S1*PPP
    fp_mul(t, p.zz, q.zz);
    fp_mul(o.zz, t, PP);
    fp_mul(t, p.zzz, q.zzz);
    fp_mul(o.zzz, t, PPP);
    r = o;
}

// canonical coordinates, for points that leave the device
template <class F>
__device__ __forceinline__ void xyzz_canon(xyzz_t<F>& r, const xyzz_t<F>& p) {
    fp_canon(r.x, p.x);
    fp_canon(r.y, p.y);
    fp_canon(r.zz, p.zz);
    fp_canon(r.zzz, p.zzz);
}

template <class F>
__device__ __forceinline__ void xyzz_neg(xyzz_t<F>& r, const xyzz_t<F>& p) {
    r = p;
    fp_neg(r.y, p.y);
}

// Jacobian (X, Y, Z) with Z = ZZ:  X = x*ZZ^2 = X'*ZZ, Y = y*ZZ^3 = Y'*ZZZ
// (Pointxyzz::to_Jacobian, point.cpp:2109-2114)
template <class F>
__device__ __forceinline__ void xyzz_to_jac(jac_t<F>& r, const xyzz_t<F>& p) {
    fp_mul(r.x, p.x, p.zz);
    fp_mul(r.y, p.y, p.zzz);
    r.z = p.zz;
}

}  // namespace bls
