// Four-lane XYZZ arithmetic for the latency-bound stages (partial merges,
// bucket reduction trees).
//
// The four lanes of a quad (lane & 3) each own one coordinate of every point
// they handle: lane 0 X, lane 1 Y, lane 2 ZZ, lane 3 ZZZ, which is also the
// field order of xyzz in memory, so a quad loads or stores a point as four
// adjacent 48-byte pieces.  An addition runs as four levels of one Fp
// product per lane instead of fourteen products in one lane; operands move
// between the lanes of a quad with DPP quad_perm moves.  Every formula,
// product and exceptional case is the one of xyzz_add / xyzz_dbl_inl in
// ec.hip.h, and Fp results are canonical, so both forms give identical
// coordinates.
//
// All four lanes of a quad must be active whenever a quad operation runs
// (control flow inside is uniform per quad).
#pragma once
#include "ec.hip.h"

namespace bls {

// quad_perm control: lane i of each quad reads lane s_i of the same quad
constexpr int quad_perm(int s0, int s1, int s2, int s3) { return s0 | (s1 << 2) | (s2 << 4) | (s3 << 6); }
constexpr int kQuadSwap = quad_perm(2, 3, 0, 1);

template <int CTRL>
__device__ __forceinline__ uint32_t qmov(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ void fp_qmov(fp& r, const fp& a) {
#pragma unroll
    for (int k = 0; k < 12; k++) r.l[k] = qmov<CTRL>(a.l[k]);
}

// value held by lane src of the quad, on every lane of the quad
template <int SRC>
__device__ __forceinline__ void fp_qbcast(fp& r, const fp& a) {
    fp_qmov<quad_perm(SRC, SRC, SRC, SRC)>(r, a);
}

template <int SRC>
__device__ __forceinline__ uint32_t u32_qbcast(uint32_t v) {
    return qmov<quad_perm(SRC, SRC, SRC, SRC)>(v);
}

__device__ __forceinline__ void fp_sel(fp& r, bool c, const fp& a, const fp& b) {
#pragma unroll
    for (int k = 0; k < 12; k++) r.l[k] = c ? a.l[k] : b.l[k];
}

template <int CTRL>
__device__ __forceinline__ void fp_qmov(fp2& r, const fp2& a) {
    fp_qmov<CTRL>(r.c0, a.c0);
    fp_qmov<CTRL>(r.c1, a.c1);
}

template <int SRC>
__device__ __forceinline__ void fp_qbcast(fp2& r, const fp2& a) {
    fp_qmov<quad_perm(SRC, SRC, SRC, SRC)>(r, a);
}

__device__ __forceinline__ void fp_sel(fp2& r, bool c, const fp2& a, const fp2& b) {
    fp_sel(r.c0, c, a.c0, b.c0);
    fp_sel(r.c1, c, a.c1, b.c1);
}

__device__ __forceinline__ uint32_t quad_lane() { return __lane_id() & 3u; }

// this lane's coordinate of point p (X, Y, ZZ or ZZZ by quad lane)
__device__ __forceinline__ void quad_load(fp& r, const xyzz* __restrict__ p) {
    const uint4* s = reinterpret_cast<const uint4*>(p) + 3 * quad_lane();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 v = s[k];
        r.l[4 * k] = v.x; r.l[4 * k + 1] = v.y; r.l[4 * k + 2] = v.z; r.l[4 * k + 3] = v.w;
    }
}

__device__ __forceinline__ void quad_store(xyzz* __restrict__ p, const fp& r) {
    uint4* d = reinterpret_cast<uint4*>(p) + 3 * quad_lane();
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = make_uint4(r.l[4 * k], r.l[4 * k + 1], r.l[4 * k + 2], r.l[4 * k + 3]);
}

// G2: the lane's Fp2 coordinate, 96 bytes
__device__ __forceinline__ void quad_load(fp2& r, const xyzz2* __restrict__ p) {
    const uint4* s = reinterpret_cast<const uint4*>(p) + 6 * quad_lane();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 v = s[k], w = s[k + 3];
        r.c0.l[4 * k] = v.x; r.c0.l[4 * k + 1] = v.y; r.c0.l[4 * k + 2] = v.z; r.c0.l[4 * k + 3] = v.w;
        r.c1.l[4 * k] = w.x; r.c1.l[4 * k + 1] = w.y; r.c1.l[4 * k + 2] = w.z; r.c1.l[4 * k + 3] = w.w;
    }
}

__device__ __forceinline__ void quad_store(xyzz2* __restrict__ p, const fp2& r) {
    uint4* d = reinterpret_cast<uint4*>(p) + 6 * quad_lane();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        d[k] = make_uint4(r.c0.l[4 * k], r.c0.l[4 * k + 1], r.c0.l[4 * k + 2], r.c0.l[4 * k + 3]);
        d[k + 3] = make_uint4(r.c1.l[4 * k], r.c1.l[4 * k + 1], r.c1.l[4 * k + 2], r.c1.l[4 * k + 3]);
    }
}

template <class F>
__device__ __forceinline__ bool quad_is_inf(const F& a) {
    return u32_qbcast<2>(fp_is_zero(a) ? 1u : 0u) != 0;
}

// r = p + q, add-2008-s with the unified doubling of xyzz_add
template <class F>
__device__ __forceinline__ void quad_add(F& r, const F& p, const F& q) {
    const uint32_t ln = quad_lane();
    const bool pinf = quad_is_inf(p), qinf = quad_is_inf(q);
    if (qinf) { r = p; return; }
    if (pinf) { r = q; return; }
    // level 1: U1 = X1*ZZ2, S1 = Y1*ZZZ2, U2 = ZZ1*X2, S2 = ZZZ1*Y2
    F b, m1;
    fp_qmov<kQuadSwap>(b, q);
    fp_mul(m1, p, b);
    F d, diff;
    fp_qmov<kQuadSwap>(d, m1);
    fp_sub(diff, d, m1);                       // lane 0: P = U2 - U1, lane 1: R = S2 - S1
    const bool pz = u32_qbcast<0>(fp_is_zero(diff) ? 1u : 0u) != 0;
    const bool rz = u32_qbcast<1>(fp_is_zero(diff) ? 1u : 0u) != 0;
    if (pz && !rz) { fp_zero(r); return; }
    const bool dbl = pz;
    F U1;
    fp_qbcast<0>(U1, m1);
    if (dbl) {
        // P' = 2*S1 (lane 0), R' = 3*U1^2 (lane 1)
        F S1, t, t2;
        fp_qbcast<1>(S1, m1);
        fp_sqr(t, U1);
        fp_dbl(t2, t);
        fp_add(t2, t2, t);
        fp_dbl(S1, S1);
        fp_sel(diff, ln == 0, S1, diff);
        fp_sel(diff, ln == 1, t2, diff);
    }
    // level 2: PP = P^2, RR = R^2, ZZ1*ZZ2, ZZZ1*ZZZ2
    F a2, b2, m2;
    fp_sel(a2, ln < 2, diff, p);
    fp_sel(b2, ln < 2, diff, q);
    fp_mul(m2, a2, b2);
    // level 3: PPP = P*PP, Q = U1*PP, ZZ3 = ZZ1*ZZ2*PP
    F PP, a3, m3;
    fp_qbcast<0>(PP, m2);
    fp_sel(a3, ln == 1, U1, m2);
    fp_sel(a3, ln == 0, diff, a3);
    fp_mul(m3, a3, PP);
    F RR, Q, PPP, X3;
    fp_qbcast<1>(RR, m2);
    fp_qbcast<1>(Q, m3);
    fp_qbcast<0>(PPP, m3);
    X3 = RR;
    if (!dbl) fp_sub(X3, X3, PPP);
    fp_sub(X3, X3, Q);
    fp_sub(X3, X3, Q);                         // X3 = R^2 - PPP - 2Q
    // level 4: S1*PPP, R*(Q - X3), ZZZ3 = ZZZ1*ZZZ2*PPP
    F S1, R, QX, a4, b4, m4;
    fp_qbcast<1>(S1, m1);
    fp_qbcast<1>(R, diff);
    fp_sub(QX, Q, X3);
    fp_sel(a4, ln == 0, S1, m2);
    fp_sel(a4, ln == 1, QX, a4);
    fp_sel(b4, ln == 1, R, PPP);
    fp_mul(m4, a4, b4);
    F y1, y0, Y3;
    fp_qbcast<1>(y1, m4);
    fp_qbcast<0>(y0, m4);
    fp_sub(Y3, y1, y0);                        // Y3 = R(Q - X3) - S1*PPP
    fp_sel(r, ln == 3, m4, m3);
    fp_sel(r, ln == 1, Y3, r);
    fp_sel(r, ln == 0, X3, r);
}

// r = 2p, dbl-2008-s-1 as xyzz_dbl_inl
template <class F>
__device__ __forceinline__ void quad_dbl(F& r, const F& p) {
    const uint32_t ln = quad_lane();
    if (quad_is_inf(p)) { r = p; return; }
    // level 1: XX = X^2 (lane 0), V = U^2 with U = 2Y (lane 1)
    F U, a1, m1;
    fp_dbl(U, p);                              // lane 1: U = 2Y
    fp_sel(a1, ln == 1, U, p);
    fp_sqr(m1, a1);
    // level 2: S = X*V, W = U*V, ZZ3 = ZZ*V, M^2 with M = 3*XX
    F V, XX, M, t, a2, b2, m2;
    fp_qbcast<1>(V, m1);
    fp_qbcast<0>(XX, m1);
    fp_dbl(t, XX);
    fp_add(M, XX, t);
    fp_sel(a2, ln == 1, U, p);
    fp_sel(a2, ln == 3, M, a2);
    fp_sel(b2, ln == 3, M, V);
    fp_mul(m2, a2, b2);
    F S, MM, X3;
    fp_qbcast<0>(S, m2);
    fp_qbcast<3>(MM, m2);
    fp_dbl(t, S);
    fp_sub(X3, MM, t);                         // X3 = M^2 - 2S
    // level 3: M*(S - X3) (lane 0), W*Y (lane 1), W*ZZZ (lane 3)
    F W, SX, a3, b3, m3;
    fp_qbcast<1>(W, m2);
    fp_sub(SX, S, X3);
    fp_sel(a3, ln == 0, SX, W);
    fp_sel(a3, ln == 1, m2, a3);
    fp_sel(b3, ln == 0, M, p);
    fp_mul(m3, a3, b3);
    F y0, y1, Y3;
    fp_qbcast<0>(y0, m3);
    fp_qbcast<1>(y1, m3);
    fp_sub(Y3, y0, y1);                        // Y3 = M(S - X3) - W*Y
    fp_sel(r, ln == 2, m2, m3);
    fp_sel(r, ln == 1, Y3, r);
    fp_sel(r, ln == 0, X3, r);
}

}  // namespace bls
