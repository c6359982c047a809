/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see blst_oracle.h).
 *
 * CPU restatement of category-blst 1.0.0's G1 MSM path.  Each function names
 * the reference file:line it follows.  This file is compiled into
 * oracle/liboracle.so, which only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline load; the product never does.
 *
 * Pinning (what checks this file independently of itself) is listed in
 * tests/test_oracle.py: Python big-integer field arithmetic, the textbook
 * affine group law, the IETF G1 generator encoding, hash-to-curve G1 vectors
 * and the reference fuzz corpus replays.
 */
#include "blst_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

/* blst/scalar/fields/Fp.h:168-192: modulus, -p^-1 mod 2^64, R mod p, R^2 mod p */
static const u64 P[6] = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
                         0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
static const u64 N0 = 0x89f3fffcfffcfffdULL;
static const u64 ONE[6] = {0x760900000002fffdULL, 0xebf4000bc40c0002ULL, 0x5f48985753c758baULL,
                           0x77ce585370525745ULL, 0x5c071a97a256ec6dULL, 0x15f65ec3fa80e493ULL};
static const u64 RR[6] = {0xf4df1f341c341746ULL, 0x0a76e6a609d104f1ULL, 0x8de5476c4c95b6d5ULL,
                          0x67eb88a9939d83c0ULL, 0x9a793e85b519952dULL, 0x11988fe592cae3aaULL};
/* blst/scalar/fields/Fr.h:127-129: group order r */
static const u64 RORD[4] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL,
                            0x73eda753299d7d48ULL};
/* G1 generator, Montgomery form (x from point.cpp:69, y its even root) */
static const u64 GX[6] = {0x5cb38790fd530c16ULL, 0x7817fc679976fff5ULL, 0x154f95c7143ba1c1ULL,
                          0xf0ae6acdf3d0e747ULL, 0xedce6ecc21dbf440ULL, 0x120177419e0bfb75ULL};
static const u64 GY[6] = {0xbaac93d50ce72271ULL, 0x8c22631a7918fd8eULL, 0xdd595f13570725ceULL,
                          0x51ac582950405194ULL, 0x0e1c8c3fad0059c0ULL, 0x0bbc3efc5008a26aULL};

/* ------------------------------------------------------------------ Fp */

static int fp_is_zero(const u64 a[6]) {
    return (a[0] | a[1] | a[2] | a[3] | a[4] | a[5]) == 0;
}
static int fp_eq(const u64 a[6], const u64 b[6]) {
    u64 d = 0;
    for (int i = 0; i < 6; i++) d |= a[i] ^ b[i];
    return d == 0;
}
static void fp_copy(u64 r[6], const u64 a[6]) { memcpy(r, a, 48); }

/* t (7 limbs, t < 2p) -> t mod p.  Final step of cc_mul_mont_n,
 * blst/backend/backend_portable.cpp:71-84. */
static void fp_final_sub(u64 r[6], const u64 t[7]) {
    u64 d[6], borrow = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)t[j] - P[j] - borrow;
        d[j] = (u64)s;
        borrow = (u64)(s >> 64) & 1;
    }
    int keep = t[6] < borrow; /* t < p */
    for (int j = 0; j < 6; j++) r[j] = keep ? t[j] : d[j];
}

/* CIOS Montgomery product a*b*2^-384 mod p,
 * blst/backend/backend_portable.cpp:19-86 (cc_mul_mont_n<6>). */
void or_fp_mul(u64 r[6], const u64 a[6], const u64 b[6]) {
    u64 t[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 6; i++) {
        u128 s;
        u64 c = 0, top;
        for (int j = 0; j < 6; j++) {
            s = (u128)a[j] * b[i] + t[j] + c;
            t[j] = (u64)s;
            c = (u64)(s >> 64);
        }
        s = (u128)t[6] + c;
        t[6] = (u64)s;
        top = (u64)(s >> 64);
        u64 m = t[0] * N0;
        s = (u128)m * P[0] + t[0];
        c = (u64)(s >> 64);
        for (int j = 1; j < 6; j++) {
            s = (u128)m * P[j] + t[j] + c;
            t[j - 1] = (u64)s;
            c = (u64)(s >> 64);
        }
        s = (u128)t[6] + c;
        t[5] = (u64)s;
        t[6] = top + (u64)(s >> 64);
    }
    fp_final_sub(r, t);
}

/* backend_portable.cpp:96-125 (cc_add_mod_n) */
void or_fp_add(u64 r[6], const u64 a[6], const u64 b[6]) {
    u64 t[7], c = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)a[j] + b[j] + c;
        t[j] = (u64)s;
        c = (u64)(s >> 64);
    }
    t[6] = c;
    fp_final_sub(r, t);
}

/* backend_portable.cpp:142-165 (cc_sub_mod_n) */
void or_fp_sub(u64 r[6], const u64 a[6], const u64 b[6]) {
    u64 d[6], borrow = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)a[j] - b[j] - borrow;
        d[j] = (u64)s;
        borrow = (u64)(s >> 64) & 1;
    }
    u64 mask = 0 - borrow, c = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)d[j] + (P[j] & mask) + c;
        r[j] = (u64)s;
        c = (u64)(s >> 64);
    }
}

/* Fp::cneg with neg=1 (Fp.h:353-361): 0 stays 0 */
void or_fp_neg(u64 r[6], const u64 a[6]) {
    if (fp_is_zero(a)) {
        memset(r, 0, 48);
        return;
    }
    u64 borrow = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)P[j] - a[j] - borrow;
        r[j] = (u64)s;
        borrow = (u64)(s >> 64) & 1;
    }
}

static void fp_cneg(u64 r[6], const u64 a[6], int neg) {
    if (neg)
        or_fp_neg(r, a);
    else
        fp_copy(r, a);
}

static void fp_mul_by_3(u64 r[6], const u64 a[6]) {
    u64 t[6];
    or_fp_add(t, a, a);
    or_fp_add(r, t, a);
}
static void fp_mul_by_8(u64 r[6], const u64 a[6]) {
    u64 t[6];
    or_fp_add(t, a, a);
    or_fp_add(t, t, t);
    or_fp_add(r, t, t);
}

void or_fp_to_mont(u64 r[6], const u64 a[6]) { or_fp_mul(r, a, RR); }

/* Fp::from_mont (Fp.h:332-335): multiply by plain 1 */
void or_fp_from_mont(u64 r[6], const u64 a[6]) {
    static const u64 one_plain[6] = {1, 0, 0, 0, 0, 0};
    or_fp_mul(r, a, one_plain);
}

/* Fermat inversion a^(p-2), the exponent of Fp.h:596-600 (flt_reciprocal).
 * 0 maps to 0, as the reference's reciprocal does. */
void or_fp_inv(u64 r[6], const u64 a[6]) {
    u64 e[6];
    memcpy(e, P, 48);
    e[0] -= 2;
    u64 acc[6];
    fp_copy(acc, ONE);
    for (int bit = 380; bit >= 0; bit--) {
        or_fp_mul(acc, acc, acc);
        if ((e[bit / 64] >> (bit % 64)) & 1) or_fp_mul(acc, acc, a);
    }
    fp_copy(r, acc);
}

/* Fp::sgn0_pty_mont (Fp.h:464-477) with cc_sgn0_pty_mod_n
 * (backend_portable.cpp:632-657): bit0 = parity, bit1 = (a > (p-1)/2). */
u64 or_fp_sgn0_pty_mont(const u64 a[6]) {
    u64 t[6];
    or_fp_from_mont(t, a);
    u64 ret = t[0] & 1;
    u64 dbl[7], carry = 0;
    for (int j = 0; j < 6; j++) {
        dbl[j] = (t[j] << 1) | carry;
        carry = t[j] >> 63;
    }
    dbl[6] = carry;
    u64 borrow = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)dbl[j] - P[j] - borrow;
        borrow = (u64)(s >> 64) & 1;
    }
    int ge = dbl[6] >= borrow; /* 2a >= p */
    if (ge) ret |= 2;
    return ret;
}

/* --------------------------------------------------------------- points */

typedef struct { u64 x[6], y[6], z[6]; } jac_t;
typedef struct { u64 x[6], y[6]; } aff_t;
/* XYZZ: x = X/ZZ, y = Y/ZZZ; infinity when ZZ == ZZZ == 0 (point.h:347-384) */
typedef struct { u64 x[6], y[6], zzz[6], zz[6]; } xyzz_t;

static int aff_is_inf(const aff_t *a) { return fp_is_zero(a->x) && fp_is_zero(a->y); }
static int xyzz_is_inf(const xyzz_t *p) { return fp_is_zero(p->zz) && fp_is_zero(p->zzz); }

/* Point::doubling, dbl-2009-l (point.cpp:1898-1925) */
static void jac_double(jac_t *p3, const jac_t *p1) {
    u64 A[6], B[6], C[6], t[6];
    or_fp_mul(A, p1->x, p1->x);
    or_fp_mul(B, p1->y, p1->y);
    or_fp_mul(C, B, B);
    or_fp_add(B, B, p1->x);
    or_fp_mul(B, B, B);
    or_fp_sub(B, B, A);
    or_fp_sub(B, B, C);
    or_fp_add(B, B, B);          /* D */
    fp_mul_by_3(A, A);           /* E */
    jac_t o;
    or_fp_mul(o.x, A, A);
    or_fp_sub(o.x, o.x, B);
    or_fp_sub(o.x, o.x, B);      /* X3 = F - 2D */
    or_fp_add(t, p1->z, p1->z);
    or_fp_mul(o.z, t, p1->y);    /* Z3 = 2 Y1 Z1 */
    fp_mul_by_8(C, C);
    or_fp_sub(o.y, B, o.x);
    or_fp_mul(o.y, o.y, A);
    or_fp_sub(o.y, o.y, C);      /* Y3 = E (D - X3) - 8C */
    *p3 = o;
}

/* Point::dadd (point.cpp:1623-1681): add with the doubling and infinity
 * cases resolved; restated with explicit branches. */
static void jac_add(jac_t *out, const jac_t *p1, const jac_t *p2) {
    if (fp_is_zero(p1->z)) { *out = *p2; return; }
    if (fp_is_zero(p2->z)) { *out = *p1; return; }
    u64 z1z1[6], z2z2[6], u1[6], u2[6], s1[6], s2[6], t[6], H[6], R[6];
    or_fp_mul(z1z1, p1->z, p1->z);
    or_fp_mul(z2z2, p2->z, p2->z);
    or_fp_mul(u1, p1->x, z2z2);
    or_fp_mul(u2, p2->x, z1z1);
    or_fp_mul(t, p2->z, z2z2);
    or_fp_mul(s1, p1->y, t);
    or_fp_mul(t, p1->z, z1z1);
    or_fp_mul(s2, p2->y, t);
    or_fp_sub(H, u2, u1);
    or_fp_sub(R, s2, s1);
    if (fp_is_zero(H)) {
        if (fp_is_zero(R)) { jac_double(out, p1); return; }
        memset(out, 0, sizeof(*out));   /* P + (-P) */
        return;
    }
    u64 HH[6], HHH[6], sx[6], V[6];
    jac_t o;
    or_fp_mul(o.z, p1->z, p2->z);
    or_fp_mul(o.z, o.z, H);            /* Z3 = H Z1 Z2 */
    or_fp_mul(HH, H, H);
    or_fp_mul(HHH, HH, H);
    or_fp_add(sx, u1, u2);
    or_fp_mul(t, HH, sx);
    or_fp_mul(o.x, R, R);
    or_fp_sub(o.x, o.x, t);            /* X3 = R^2 - H^2 (U1+U2) */
    or_fp_mul(V, HH, u1);
    or_fp_sub(V, V, o.x);
    or_fp_mul(V, V, R);
    or_fp_mul(t, HHH, s1);
    or_fp_sub(o.y, V, t);              /* Y3 = R (H^2 U1 - X3) - H^3 S1 */
    *out = o;
}

/* Point::from_affine (point.cpp:665-670): all-zero affine -> Z = 0 */
static void jac_from_aff(jac_t *out, const aff_t *a) {
    fp_copy(out->x, a->x);
    fp_copy(out->y, a->y);
    if (aff_is_inf(a))
        memset(out->z, 0, 48);
    else
        fp_copy(out->z, ONE);
}

/* Point::from_jacobian (point.cpp:628-643); infinity -> all zero affine */
static void jac_to_aff(aff_t *out, const jac_t *p) {
    if (fp_is_zero(p->z)) { memset(out, 0, sizeof(*out)); return; }
    u64 zi[6], zi2[6], zi3[6];
    or_fp_inv(zi, p->z);
    or_fp_mul(zi2, zi, zi);
    or_fp_mul(zi3, zi2, zi);
    or_fp_mul(out->x, p->x, zi2);
    or_fp_mul(out->y, p->y, zi3);
}

/* Pointxyzz::dadd(xyzz, affine, subtract): madd-2008-s / mdbl-2008-s-1
 * (point.cpp:2027-2091) */
static void xyzz_add_aff(xyzz_t *p3, const xyzz_t *p1, const aff_t *p2, int subtract) {
    if (aff_is_inf(p2)) { *p3 = *p1; return; }
    if (xyzz_is_inf(p1)) {
        fp_copy(p3->x, p2->x);
        fp_copy(p3->y, p2->y);
        fp_cneg(p3->zzz, ONE, subtract);
        fp_copy(p3->zz, ONE);
        return;
    }
    u64 Pv[6], R[6];
    or_fp_mul(Pv, p2->x, p1->zz);
    or_fp_mul(R, p2->y, p1->zzz);
    fp_cneg(R, R, subtract);
    or_fp_sub(Pv, Pv, p1->x);
    or_fp_sub(R, R, p1->y);
    xyzz_t o;
    if (!fp_is_zero(Pv)) {
        u64 PP[6], PPP[6], Q[6], t[6];
        or_fp_mul(PP, Pv, Pv);
        or_fp_mul(PPP, PP, Pv);
        or_fp_mul(Q, p1->x, PP);
        or_fp_mul(o.x, R, R);
        or_fp_add(t, Q, Q);
        or_fp_sub(o.x, o.x, PPP);
        or_fp_sub(o.x, o.x, t);
        or_fp_sub(Q, Q, o.x);
        or_fp_mul(Q, Q, R);
        or_fp_mul(o.y, p1->y, PPP);
        or_fp_sub(o.y, Q, o.y);
        or_fp_mul(o.zz, p1->zz, PP);
        or_fp_mul(o.zzz, p1->zzz, PPP);
    } else if (fp_is_zero(R)) {
        u64 U[6], S[6], M[6];
        or_fp_add(U, p2->y, p2->y);
        or_fp_mul(o.zz, U, U);
        or_fp_mul(o.zzz, o.zz, U);
        or_fp_mul(S, p2->x, o.zz);
        or_fp_mul(M, p2->x, p2->x);
        fp_mul_by_3(M, M);
        or_fp_mul(o.x, M, M);
        or_fp_add(U, S, S);
        or_fp_sub(o.x, o.x, U);
        or_fp_mul(o.y, o.zzz, p2->y);
        or_fp_sub(S, S, o.x);
        or_fp_mul(S, S, M);
        or_fp_sub(o.y, S, o.y);
        fp_cneg(o.zzz, o.zzz, subtract);
    } else {
        memset(&o, 0, sizeof(o));
    }
    *p3 = o;
}

/* Pointxyzz::dadd(xyzz, xyzz): add-2008-s / dbl-2008-s-1 (point.cpp:1939-2001) */
static void xyzz_add(xyzz_t *p3, const xyzz_t *p1, const xyzz_t *p2) {
    if (xyzz_is_inf(p2)) { *p3 = *p1; return; }
    if (xyzz_is_inf(p1)) { *p3 = *p2; return; }
    u64 U[6], S[6], Pv[6], R[6];
    or_fp_mul(U, p1->x, p2->zz);
    or_fp_mul(S, p1->y, p2->zzz);
    or_fp_mul(Pv, p2->x, p1->zz);
    or_fp_mul(R, p2->y, p1->zzz);
    or_fp_sub(Pv, Pv, U);
    or_fp_sub(R, R, S);
    xyzz_t o;
    if (!fp_is_zero(Pv)) {
        u64 PP[6], PPP[6], Q[6], t[6];
        or_fp_mul(PP, Pv, Pv);
        or_fp_mul(PPP, PP, Pv);
        or_fp_mul(Q, U, PP);
        or_fp_mul(o.x, R, R);
        or_fp_add(t, Q, Q);
        or_fp_sub(o.x, o.x, PPP);
        or_fp_sub(o.x, o.x, t);
        or_fp_sub(Q, Q, o.x);
        or_fp_mul(Q, Q, R);
        or_fp_mul(o.y, S, PPP);
        or_fp_sub(o.y, Q, o.y);
        or_fp_mul(o.zz, p1->zz, p2->zz);
        or_fp_mul(o.zzz, p1->zzz, p2->zzz);
        or_fp_mul(o.zz, o.zz, PP);
        or_fp_mul(o.zzz, o.zzz, PPP);
    } else if (fp_is_zero(R)) {
        u64 V[6], W[6], M[6], t[6];
        or_fp_add(U, p1->y, p1->y);
        or_fp_mul(V, U, U);
        or_fp_mul(W, V, U);
        or_fp_mul(S, p1->x, V);
        or_fp_mul(M, p1->x, p1->x);
        fp_mul_by_3(M, M);
        or_fp_mul(o.x, M, M);
        or_fp_add(t, S, S);
        or_fp_sub(o.x, o.x, t);
        or_fp_mul(o.y, W, p1->y);
        or_fp_sub(S, S, o.x);
        or_fp_mul(S, S, M);
        or_fp_sub(o.y, S, o.y);
        or_fp_mul(o.zz, p1->zz, V);
        or_fp_mul(o.zzz, p1->zzz, W);
    } else {
        memset(&o, 0, sizeof(o));
    }
    *p3 = o;
}

/* Pointxyzz::to_Jacobian (point.cpp:2109-2114) */
static void xyzz_to_jac(jac_t *out, const xyzz_t *in) {
    jac_t o;
    or_fp_mul(o.x, in->x, in->zz);
    or_fp_mul(o.y, in->y, in->zzz);
    fp_copy(o.z, in->zz);
    *out = o;
}

void or_p1_generator(u64 out[12]) {
    memcpy(out, GX, 48);
    memcpy(out + 6, GY, 48);
}
void or_p1_from_affine(u64 out[18], const u64 a[12]) { jac_from_aff((jac_t *)out, (const aff_t *)a); }
void or_p1_add(u64 out[18], const u64 a[18], const u64 b[18]) {
    jac_add((jac_t *)out, (const jac_t *)a, (const jac_t *)b);
}
void or_p1_double(u64 out[18], const u64 a[18]) { jac_double((jac_t *)out, (const jac_t *)a); }
void or_p1_to_affine(u64 out[12], const u64 p[18]) { jac_to_aff((aff_t *)out, (const jac_t *)p); }

/* Point::is_equal semantics: projective comparison, infinities equal */
int or_p1_is_equal(const u64 a[18], const u64 b[18]) {
    const jac_t *p = (const jac_t *)a, *q = (const jac_t *)b;
    int ai = fp_is_zero(p->z), bi = fp_is_zero(q->z);
    if (ai || bi) return ai && bi;
    u64 z1z1[6], z2z2[6], l[6], r[6], t[6];
    or_fp_mul(z1z1, p->z, p->z);
    or_fp_mul(z2z2, q->z, q->z);
    or_fp_mul(l, p->x, z2z2);
    or_fp_mul(r, q->x, z1z1);
    if (!fp_eq(l, r)) return 0;
    or_fp_mul(t, z2z2, q->z);
    or_fp_mul(l, p->y, t);
    or_fp_mul(t, z1z1, p->z);
    or_fp_mul(r, q->y, t);
    return fp_eq(l, r);
}

/* y^2 == x^3 + 4 (infinity counts as on the curve) */
int or_p1_affine_on_curve(const u64 a[12]) {
    const aff_t *p = (const aff_t *)a;
    if (aff_is_inf(p)) return 1;
    u64 l[6], r[6], four[6] = {4, 0, 0, 0, 0, 0}, b[6];
    or_fp_to_mont(b, four);
    or_fp_mul(l, p->y, p->y);
    or_fp_mul(r, p->x, p->x);
    or_fp_mul(r, r, p->x);
    or_fp_add(r, r, b);
    return fp_eq(l, r);
}

/* Point::compress (point.cpp:892-902) -> compress_BE (:874-886) ->
 * compress_affine_coordinates_BE (:817-842), be_bytes_from_limbs (bytes.h:42-51) */
void or_p1_compress(uint8_t out[48], const u64 p[18]) {
    const jac_t *in = (const jac_t *)p;
    if (fp_is_zero(in->z)) {
        memset(out, 0, 48);
        out[0] = 0xc0;
        return;
    }
    aff_t a;
    if (fp_eq(in->z, ONE)) {
        fp_copy(a.x, in->x);
        fp_copy(a.y, in->y);
    } else {
        jac_to_aff(&a, in);
    }
    u64 x[6];
    or_fp_from_mont(x, a.x);
    for (int n = 47, k = 0; n >= 0; n--, k++) out[k] = (uint8_t)(x[n / 8] >> (8 * (n % 8)));
    u64 sign = or_fp_sgn0_pty_mont(a.y);
    out[0] |= (uint8_t)(0x80 | ((sign & 2) << 4));
}

/* points_to_affine (multi_scalar.cpp:160-238): Montgomery's trick, one
 * inversion per call; infinity stays all-zero. */
void or_p1s_to_affine(u64 *out_affine, const u64 *jacobian, size_t n) {
    if (n == 0) return;
    const jac_t *in = (const jac_t *)jacobian;
    aff_t *out = (aff_t *)out_affine;
    u64(*pre)[6] = malloc(n * 48);
    u64 acc[6];
    fp_copy(acc, ONE);
    for (size_t i = 0; i < n; i++) {
        fp_copy(pre[i], acc);
        if (!fp_is_zero(in[i].z)) or_fp_mul(acc, acc, in[i].z);
    }
    u64 inv[6];
    or_fp_inv(inv, acc);
    for (size_t i = n; i-- > 0;) {
        if (fp_is_zero(in[i].z)) {
            memset(&out[i], 0, sizeof(aff_t));
            continue;
        }
        u64 zi[6], zi2[6], zi3[6];
        or_fp_mul(zi, inv, pre[i]);
        or_fp_mul(inv, inv, in[i].z);
        or_fp_mul(zi2, zi, zi);
        or_fp_mul(zi3, zi2, zi);
        or_fp_mul(out[i].x, in[i].x, zi2);
        or_fp_mul(out[i].y, in[i].y, zi3);
    }
    free(pre);
}

/* Test inputs for batch normalisation: affine (x, y) to the Jacobian
 * representative (x l^2, y l^3, l) with l = lambda[i] (Montgomery form);
 * an all-zero affine point gives Z = 0. */
void or_p1s_from_affine_scaled(u64 *jacobian, const u64 *affine, const u64 *lambda, size_t n) {
    for (size_t i = 0; i < n; i++) {
        const aff_t *a = (const aff_t *)(affine + 12 * i);
        jac_t *j = (jac_t *)(jacobian + 18 * i);
        const u64 *l = lambda + 6 * i;
        if (fp_is_zero(a->x) && fp_is_zero(a->y)) {
            memset(j, 0, sizeof(*j));
            continue;
        }
        u64 l2[6], l3[6];
        or_fp_mul(l2, l, l);
        or_fp_mul(l3, l2, l);
        or_fp_mul(j->x, a->x, l2);
        or_fp_mul(j->y, a->y, l3);
        fp_copy(j->z, l);
    }
}

/* Variable-time double-and-add over the low nbits of the scalar (the value
 * mult_vartime computes, point.cpp:1546-1565). */
void or_p1_mult(u64 out[18], const u64 p[18], const uint8_t *scalar, size_t nbits) {
    jac_t acc;
    memset(&acc, 0, sizeof(acc));
    const jac_t *base = (const jac_t *)p;
    for (size_t bit = nbits; bit-- > 0;) {
        jac_double(&acc, &acc);
        if ((scalar[bit / 8] >> (bit % 8)) & 1) jac_add(&acc, &acc, base);
    }
    *(jac_t *)out = acc;
}

/* ------------------------------------------------------------- Pippenger */

/* pippenger_window_size (multi_scalar.h:329-348) */
size_t or_pippenger_window_size(size_t npoints) {
    size_t wbits;
    for (wbits = 0; npoints >>= 1; wbits++) {}
    if (wbits > 12)
        wbits -= 3;
    else if (wbits > 8)
        wbits -= 2;
    else if (wbits > 4)
        wbits -= 1;
    else
        return wbits ? 2 : 1;
    return wbits < 24 ? wbits : 24;
}

/* msm_mult_pippenger_scratch_sizeof_<Pointxyzz<Fp>> (multi_scalar.h:460-469):
 * 192 bytes per bucket, 2^(window-1) buckets, 0 for no points */
size_t or_pippenger_scratch_sizeof(size_t npoints) {
    if (npoints == 0) return 0;
    return (size_t)192 << (or_pippenger_window_size(npoints) - 1);
}

/* msm_mult_pippenger_window_scratch_sizeof_ (multi_scalar.h:471-479) with
 * valid_pippenger_window = [1, 24] */
size_t or_pippenger_window_scratch_sizeof(size_t window) {
    if (window < 1 || window > 24) return 0;
    return (size_t)192 << (window - 1);
}

/* ------------------------------------------------------------ blst_fr */

/* Fr::to_scaler (blst/scalar/fields/Fr.h:445-450) = Fr::from_mont
 * (Fr.h:634-637, from_mont_256 of blst/backend/backend.h:91-98): the
 * Montgomery form x of a scalar (R = 2^256 mod r) to k = x * 2^-256 mod r,
 * as four rounds of limb-wise Montgomery reduction by r and one conditional
 * subtraction, written out little-endian.  Accepts any 256-bit x: the
 * reduction leaves t <= r before the subtraction. */
void or_fr_to_scalar(uint8_t out[32], const u64 fr[4]) {
    const u64 r0 = 0xfffffffeffffffffULL;   /* -r^-1 mod 2^64 */
    u64 t[5] = {fr[0], fr[1], fr[2], fr[3], 0};
    for (int k = 0; k < 4; k++) {
        const u64 m = t[0] * r0;
        u128 acc = (u128)m * RORD[0] + t[0];
        acc >>= 64;
        for (int j = 1; j < 4; j++) {
            acc += (u128)m * RORD[j] + t[j];
            t[j - 1] = (u64)acc;
            acc >>= 64;
        }
        acc += t[4];
        t[3] = (u64)acc;
        t[4] = (u64)(acc >> 64);
    }
    u64 d[4], borrow = 0;
    for (int j = 0; j < 4; j++) {
        const u128 v = (u128)t[j] - RORD[j] - borrow;
        d[j] = (u64)v;
        borrow = (u64)(v >> 64) & 1;
    }
    const int ge = t[4] != 0 || borrow == 0;
    for (int j = 0; j < 4; j++) {
        const u64 v = ge ? d[j] : t[j];
        for (int b = 0; b < 8; b++) out[8 * j + b] = (uint8_t)(v >> (8 * b));
    }
}

/* c_api_scalars_from_fr (blst/c_api.cpp:315-333) over n contiguous values */
void or_frs_to_scalars(uint8_t *out, const u64 *fr, size_t n) {
    for (size_t i = 0; i < n; i++) or_fr_to_scalar(out + 32 * i, fr + 4 * i);
}

/* The scratch-size queries of the blst_fr entry points for either group:
 * the element types of one group's layouts
 * (blst/scalar/curve/multi_scalar_batch_affine.cpp:16-36; Point<Field> is
 * Jacobian (X, Y, Z), Pointxyzz<Field> four coordinates) and the largest size
 * the planner gives to batch-affine (c_api.cpp:700-719: 2^18 points for G1,
 * 2^14 for G2). */
typedef struct { u64 x[6], y[6]; uint32_t generation, queued_epoch; } ba_affine_bucket_t;
typedef struct { u64 point[18]; uint32_t generation; } ba_projective_bucket_t;
typedef struct { u64 x[12], y[12]; uint32_t generation, queued_epoch; } ba2_affine_bucket_t;
typedef struct { u64 point[36]; uint32_t generation; } ba2_projective_bucket_t;
typedef struct { uint32_t point_index, bucket_index, subtract; } ba_schedule_point_t;

typedef struct {
    size_t xyzz, affine, affine_align, projective, projective_align, field, measured_max;
} fr_layout_t;

static const fr_layout_t kG1 = {4 * 48, sizeof(ba_affine_bucket_t), _Alignof(ba_affine_bucket_t),
                                sizeof(ba_projective_bucket_t), _Alignof(ba_projective_bucket_t), 48,
                                (size_t)1 << 18};
static const fr_layout_t kG2 = {4 * 96, sizeof(ba2_affine_bucket_t), _Alignof(ba2_affine_bucket_t),
                                sizeof(ba2_projective_bucket_t), _Alignof(ba2_projective_bucket_t), 96,
                                (size_t)1 << 14};

/* c_api_pippenger_fr_scratch_sizeof (blst/c_api.cpp:272-299): 32 bytes per
 * canonical scalar, and for more than one point 2^(window-1) XYZZ buckets
 * (the GroupZZ of Pippenger), 0 on overflow. */
static size_t pippenger_fr_scratch(size_t npoints, const fr_layout_t *G) {
    if (npoints == 0) return 0;
    if (npoints > SIZE_MAX / 32) return 0;
    const size_t scalar_bytes = npoints * 32;
    if (npoints == 1) return scalar_bytes;
    const size_t window = or_pippenger_window_size(npoints);
    if (window == 0 || window - 1 >= 64) return 0;
    const size_t buckets = (size_t)1 << (window - 1);
    if (buckets > (SIZE_MAX - scalar_bytes) / G->xyzz) return 0;
    return scalar_bytes + buckets * G->xyzz;
}

static size_t ba_align(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

/* batch_affine_scratch_size (multi_scalar_batch_affine.cpp:149-168): the
 * five arrays in order (affine buckets, projective buckets, the schedule,
 * two field arrays of the batch), each at its element's alignment, the total
 * rounded to the limb size; window outside [1, 24] gives 0 */
static size_t batch_affine_scratch_size(size_t window, size_t batch, const fr_layout_t *G) {
    if (window < 1 || window > 24) return 0;
    const size_t B = (size_t)1 << (window - 1);
    size_t off = 0;
    off = ba_align(off, G->affine_align) + B * G->affine;
    off = ba_align(off, G->projective_align) + B * G->projective;
    off = ba_align(off, _Alignof(ba_schedule_point_t)) + batch * sizeof(ba_schedule_point_t);
    off = ba_align(off, sizeof(u64)) + batch * G->field;
    off = ba_align(off, sizeof(u64)) + batch * G->field;
    return ba_align(off, sizeof(u64));
}

/* c_api_batch_affine_fr_scratch_sizeof (blst/c_api.cpp:593-627) */
static size_t batch_affine_fr_scratch(size_t npoints, size_t window, size_t batch, const fr_layout_t *G) {
    if (npoints == 0 || npoints > 0xffffffffULL) return 0;
    if (batch != 128 && batch != 384) return 0;
    const size_t alg = batch_affine_scratch_size(window, batch, G);
    if (alg == 0) return 0;
    if (npoints > SIZE_MAX / 32) return 0;
    const size_t scalar_bytes = npoints * 32;
    if (npoints == 1) return scalar_bytes;
    if (alg > SIZE_MAX - scalar_bytes) return 0;
    return scalar_bytes + alg;
}

/* c_api_msm_fr_plan (blst/c_api.cpp:719-755): batch-affine from 2^12 points
 * to the group's measured maximum, window by size, batch 384 from 2^18
 * points; 0 = Pippenger */
static size_t msm_fr_plan(size_t n, size_t *batch, const fr_layout_t *G) {
    if (n < ((size_t)1 << 12) || n > G->measured_max) return 0;
    size_t w;
    if (n <= ((size_t)1 << 12)) w = 10;
    else if (n <= ((size_t)1 << 14)) w = 11;
    else if (n <= ((size_t)1 << 15)) w = 12;
    else if (n <= ((size_t)1 << 16)) w = 13;
    else if (n <= ((size_t)1 << 17)) w = 14;
    else w = 15;
    *batch = n >= ((size_t)1 << 18) ? 384 : 128;
    return w;
}

/* c_api_msm_fr_scratch_sizeof (blst/c_api.cpp:758-807): a capacity query */
static size_t msm_fr_scratch(size_t max_points, const fr_layout_t *G) {
    if (max_points == 0) return 0;
    const size_t pip = pippenger_fr_scratch(max_points, G);
    if (pip == 0) return 0;
    size_t required = pip, batch = 0, w = msm_fr_plan(max_points, &batch, G);
    if (w) {
        const size_t b = batch_affine_fr_scratch(max_points, w, batch, G);
        if (b == 0) return 0;
        if (b > required) required = b;
    }
    if (max_points > G->measured_max) {
        w = msm_fr_plan(G->measured_max, &batch, G);
        const size_t b = batch_affine_fr_scratch(G->measured_max, w, batch, G);
        if (b == 0) return 0;
        if (b > required) required = b;
    }
    return required;
}

size_t or_pippenger_fr_scratch_sizeof(size_t npoints) { return pippenger_fr_scratch(npoints, &kG1); }
size_t or_batch_affine_fr_scratch_sizeof(size_t npoints, size_t window, size_t batch) {
    return batch_affine_fr_scratch(npoints, window, batch, &kG1);
}
size_t or_msm_fr_scratch_sizeof(size_t max_points) { return msm_fr_scratch(max_points, &kG1); }
size_t or_p2_pippenger_fr_scratch_sizeof(size_t npoints) { return pippenger_fr_scratch(npoints, &kG2); }
size_t or_p2_batch_affine_fr_scratch_sizeof(size_t npoints, size_t window, size_t batch) {
    return batch_affine_fr_scratch(npoints, window, batch, &kG2);
}
size_t or_p2_msm_fr_scratch_sizeof(size_t max_points) { return msm_fr_scratch(max_points, &kG2); }

/* get_wval_limb (ec_mult.h:63-82): up to 25 bits starting at bit |off|,
 * reading only the bytes that hold them */
static u64 get_wval_limb(const uint8_t *d, size_t off, size_t bits) {
    size_t first = off / 8, last = (off + bits - 1) / 8;
    u64 ret = 0;
    for (size_t i = first; i <= last; i++) ret |= (u64)d[i] << (8 * (i - first));
    return ret >> (off % 8);
}

/* booth_encode (ec_mult.h:125-142) */
static u64 booth_encode(u64 wval, size_t w) {
    u64 sign = wval >> w;
    u64 mask = 0 - sign;
    u64 val = (wval + 1) >> 1;
    return (val ^ mask) - mask;
}

typedef struct {
    size_t missing_overlap, extract_begin, extract_size, encode_size;
    u64 mask;
} wex_t;

/* WindowExtract::initialize (multi_scalar.h:387-397) */
static wex_t wex_init(size_t bit0, size_t window, size_t nbits) {
    wex_t w;
    size_t end = bit0 + window < nbits ? bit0 + window : nbits;
    w.missing_overlap = bit0 == 0;
    w.extract_begin = bit0 ? bit0 - 1 : 0;
    w.extract_size = end - w.extract_begin;
    w.mask = ((u64)1 << w.extract_size) - 1;
    w.encode_size = nbits - bit0 + 1 < window ? nbits - bit0 + 1 : window;
    return w;
}

/* WindowExtract::extract (multi_scalar.h:400-404) */
static u64 wex_extract(const wex_t *w, const uint8_t *scalar) {
    u64 wval = (get_wval_limb(scalar, w->extract_begin, w->extract_size) & w->mask)
               << w->missing_overlap;
    return booth_encode(wval, w->encode_size);
}

u64 or_booth_digit(const uint8_t *scalar, size_t nbits, size_t bit0, size_t window) {
    wex_t w = wex_init(bit0, window, nbits);
    return wex_extract(&w, scalar);
}

/* point_integrate_buckets (multi_scalar.cpp:730-751): sum_d d * S_d */
static void integrate_buckets(jac_t *out, xyzz_t *buckets, size_t wbits) {
    size_t n = (size_t)1 << wbits;
    xyzz_t acc, ret;
    acc = buckets[--n];
    ret = buckets[n];
    memset(&buckets[n], 0, sizeof(xyzz_t));
    while (n--) {
        xyzz_add(&acc, &acc, &buckets[n]);
        xyzz_add(&ret, &ret, &acc);
        memset(&buckets[n], 0, sizeof(xyzz_t));
    }
    xyzz_to_jac(out, &ret);
}

/* points_tile_pippenger (multi_scalar.cpp:806-865) for one window starting
 * at bit0, with point_bucket (:755-767) */
static void tile_pippenger(jac_t *out, const aff_t *points, const uint8_t *scalars, size_t n,
                           size_t nbits, xyzz_t *buckets, size_t bit0, size_t window) {
    wex_t w = wex_init(bit0, window, nbits);
    size_t nbytes = (nbits + 7) / 8;
    u64 idx_mask = ((u64)1 << w.encode_size) - 1;
    for (size_t i = 0; i < n; i++) {
        u64 d = wex_extract(&w, scalars + i * nbytes);
        int sign = (int)((d >> w.encode_size) & 1);
        u64 idx = d & idx_mask;
        if (idx) xyzz_add_aff(&buckets[idx - 1], &buckets[idx - 1], &points[i], sign);
    }
    integrate_buckets(out, buckets, w.encode_size - 1);
}

void or_msm_tile(u64 ret[18], const u64 *points, const uint8_t *scalars, size_t npoints,
                 size_t nbits, size_t bit0, size_t window) {
    xyzz_t *buckets = calloc((size_t)1 << (window - 1), sizeof(xyzz_t));
    tile_pippenger((jac_t *)ret, (const aff_t *)points, scalars, npoints, nbits, buckets, bit0,
                   window);
    free(buckets);
}

/* msm_mult_pippenger (multi_scalar.cpp:892-925), contiguous arrays */
void or_msm_pippenger(u64 ret[18], const u64 *points, const uint8_t *scalars, size_t npoints,
                      size_t nbits, size_t window) {
    jac_t *r = (jac_t *)ret;
    memset(r, 0, sizeof(*r));
    if (npoints == 0 || nbits == 0 || nbits > 256) return;
    if (window == 0) window = or_pippenger_window_size(npoints);
    xyzz_t *buckets = calloc((size_t)1 << (window - 1), sizeof(xyzz_t));
    size_t nwin = 1 + nbits / window;
    for (size_t wi = nwin; wi-- > 0;) {
        jac_t tile;
        tile_pippenger(&tile, (const aff_t *)points, scalars, npoints, nbits, buckets,
                       wi * window, window);
        jac_add(r, r, &tile);
        if (wi)
            for (size_t k = 0; k < window; k++) jac_double(r, r);
    }
    free(buckets);
}

typedef struct {
    const aff_t *points;
    const uint8_t *scalars;
    size_t n, nbits, window, nwin;
    int tid, nthreads;
    jac_t *tiles;
} mt_job_t;

static void *mt_worker(void *arg) {
    mt_job_t *j = (mt_job_t *)arg;
    xyzz_t *buckets = calloc((size_t)1 << (j->window - 1), sizeof(xyzz_t));
    for (size_t wi = (size_t)j->tid; wi < j->nwin; wi += (size_t)j->nthreads)
        tile_pippenger(&j->tiles[wi], j->points, j->scalars, j->n, j->nbits, buckets,
                       wi * j->window, j->window);
    free(buckets);
    return NULL;
}

/* Same result as or_msm_pippenger with the windows spread over threads and
 * combined on the caller (the blst_p1s_tile_pippenger recombination of
 * test/C_api_tests.cpp:1709-1757). */
void or_msm_pippenger_mt(u64 ret[18], const u64 *points, const uint8_t *scalars, size_t npoints,
                         size_t nbits, size_t window, int threads) {
    jac_t *r = (jac_t *)ret;
    memset(r, 0, sizeof(*r));
    if (npoints == 0 || nbits == 0 || nbits > 256) return;
    if (window == 0) window = or_pippenger_window_size(npoints);
    if (threads < 1) threads = 1;
    size_t nwin = 1 + nbits / window;
    jac_t *tiles = calloc(nwin, sizeof(jac_t));
    pthread_t *th = calloc((size_t)threads, sizeof(pthread_t));
    mt_job_t *jobs = calloc((size_t)threads, sizeof(mt_job_t));
    for (int t = 0; t < threads; t++) {
        jobs[t] = (mt_job_t){(const aff_t *)points, scalars, npoints, nbits, window, nwin,
                             t, threads, tiles};
        pthread_create(&th[t], NULL, mt_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    for (size_t wi = nwin; wi-- > 0;) {
        jac_add(r, r, &tiles[wi]);
        if (wi)
            for (size_t k = 0; k < window; k++) jac_double(r, r);
    }
    free(tiles);
    free(th);
    free(jobs);
}

/* Many independent MSMs (or_msm_pippenger each) handed out to threads, the
 * largest first: the reference library is reentrant and a batch of MSMs is
 * independent calls (README.md:39-40), so this is what batched GPU launches
 * are checked against.  rets: njobs x 18 limbs. */
typedef struct {
    u64 *rets;
    const u64 *const *points;
    const uint8_t *const *scalars;
    const size_t *ns, *nbits, *order;
    size_t njobs;
    size_t next; /* shared job counter */
} batch_job_t;

static void *batch_worker(void *arg) {
    batch_job_t *b = (batch_job_t *)arg;
    for (;;) {
        size_t k = __atomic_fetch_add(&b->next, 1, __ATOMIC_RELAXED);
        if (k >= b->njobs) return NULL;
        size_t j = b->order[k];
        or_msm_pippenger(b->rets + 18 * j, b->points[j], b->scalars[j], b->ns[j], b->nbits[j], 0);
    }
}

void or_msm_batch_mt(u64 *rets, const u64 *const *points, const uint8_t *const *scalars,
                     const size_t *ns, const size_t *nbits, size_t njobs, int threads) {
    if (threads < 1) threads = 1;
    size_t *order = calloc(njobs ? njobs : 1, sizeof(size_t));
    for (size_t j = 0; j < njobs; j++) order[j] = j;
    /* largest first (insertion sort on a copy of the sizes is enough here) */
    for (size_t a = 1; a < njobs; a++) {
        size_t v = order[a], b = a;
        while (b > 0 && ns[order[b - 1]] < ns[v]) { order[b] = order[b - 1]; b--; }
        order[b] = v;
    }
    batch_job_t job = {rets, points, scalars, ns, nbits, order, njobs, 0};
    pthread_t *th = calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, batch_worker, &job);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(order);
}

/* Sum of [k_i]P_i term by term: the "expected" of C_api_tests.cpp:1488-1515
 * and adapter_msm.cpp:19-50. */
void or_msm_naive(u64 ret[18], const u64 *points, const uint8_t *scalars, size_t npoints,
                  size_t nbits) {
    jac_t acc;
    memset(&acc, 0, sizeof(acc));
    size_t nbytes = (nbits + 7) / 8;
    for (size_t i = 0; i < npoints; i++) {
        jac_t p, t;
        jac_from_aff(&p, (const aff_t *)(points + 12 * i));
        or_p1_mult((u64 *)&t, (const u64 *)&p, scalars + i * nbytes, nbits);
        jac_add(&acc, &acc, &t);
    }
    *(jac_t *)ret = acc;
}

/* --------------------------------------------------- synthetic inputs */

/* xorshift64* (test/InsecureRng.h) */
static u64 rng_next(u64 *s) {
    u64 x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 2685821657736338717ULL;
}

static int lt_r(const u64 v[4]) {
    for (int i = 3; i >= 0; i--) {
        if (v[i] != RORD[i]) return v[i] < RORD[i];
    }
    return 0;
}

/* uniform element of Fr: 4 draws, clear bit 255, reject if >= r */
static void draw_fr(u64 *state, u64 v[4]) {
    for (;;) {
        for (int i = 0; i < 4; i++) v[i] = rng_next(state);
        v[3] &= 0x7fffffffffffffffULL;
        if (lt_r(v)) return;
    }
}

static void store_le32(uint8_t *out, const u64 v[4]) {
    for (int i = 0; i < 32; i++) out[i] = (uint8_t)(v[i / 8] >> (8 * (i % 8)));
}

/* n uniform Fr scalars, 32 bytes little-endian each.  low64_pct percent of
 * them (chosen by the same stream) keep only their low 64 bits. */
void or_gen_scalars(uint8_t *out32, size_t n, uint64_t seed, int low64_pct) {
    u64 s = seed ? seed : 1;
    for (size_t i = 0; i < n; i++) {
        u64 v[4];
        draw_fr(&s, v);
        if (low64_pct > 0 && (int)(rng_next(&s) % 100) < low64_pct) v[1] = v[2] = v[3] = 0;
        store_le32(out32 + 32 * i, v);
    }
}

/* Fixed-base comb table for G: T[m][j-1] = [j * 2^(8m)] G, m < 32, j < 256 */
static aff_t *g_table = NULL;
static pthread_once_t g_table_once = PTHREAD_ONCE_INIT;

static void build_g_table(void) {
    size_t cnt = 32 * 255;
    jac_t *jt = malloc(cnt * sizeof(jac_t));
    jac_t base;
    aff_t g;
    memcpy(g.x, GX, 48);
    memcpy(g.y, GY, 48);
    jac_from_aff(&base, &g);
    for (int m = 0; m < 32; m++) {
        jac_t acc = base;
        for (int j = 1; j < 256; j++) {
            if (j > 1) jac_add(&acc, &acc, &base);
            jt[m * 255 + (j - 1)] = acc;
        }
        for (int k = 0; k < 8; k++) jac_double(&base, &base);
    }
    aff_t *t = malloc(cnt * sizeof(aff_t));
    or_p1s_to_affine((u64 *)t, (const u64 *)jt, cnt);
    free(jt);
    g_table = t;
}

/* [k]G by the comb table: one addition per nonzero scalar byte */
static void mul_g(jac_t *out, const uint8_t k[32]) {
    jac_t acc;
    memset(&acc, 0, sizeof(acc));
    for (int m = 0; m < 32; m++) {
        if (!k[m]) continue;
        jac_t q;
        jac_from_aff(&q, &g_table[m * 255 + (k[m] - 1)]);
        jac_add(&acc, &acc, &q);
    }
    *out = acc;
}

typedef struct {
    u64 *out;
    const uint8_t *keys;
    size_t begin, end;
} gen_job_t;

static void *gen_worker(void *arg) {
    gen_job_t *j = (gen_job_t *)arg;
    const size_t block = 1024;
    jac_t *tmp = malloc(block * sizeof(jac_t));
    for (size_t b = j->begin; b < j->end; b += block) {
        size_t m = j->end - b < block ? j->end - b : block;
        for (size_t i = 0; i < m; i++) mul_g(&tmp[i], j->keys + 32 * (b + i));
        or_p1s_to_affine(j->out + 12 * b, (const u64 *)tmp, m);
    }
    free(tmp);
    return NULL;
}

static void gen_from_keys(u64 *out, const uint8_t *keys, size_t n, int threads) {
    pthread_once(&g_table_once, build_g_table);
    if (threads < 1) threads = 1;
    pthread_t *th = calloc((size_t)threads, sizeof(pthread_t));
    gen_job_t *jobs = calloc((size_t)threads, sizeof(gen_job_t));
    size_t per = (n + (size_t)threads - 1) / (size_t)threads;
    for (int t = 0; t < threads; t++) {
        size_t b = per * (size_t)t, e = b + per;
        if (b > n) b = n;
        if (e > n) e = n;
        jobs[t] = (gen_job_t){out, keys, b, e};
        pthread_create(&th[t], NULL, gen_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* P_i = [k_i]G with k_i uniform in Fr from the seeded stream, affine */
void or_gen_points(u64 *out_affine, size_t n, uint64_t seed, int threads) {
    uint8_t *keys = malloc(n * 32 + 1);
    or_gen_scalars(keys, n, seed, 0);
    gen_from_keys(out_affine, keys, n, threads);
    free(keys);
}

/* KZG setup [tau^i]G, i < n, tau uniform in Fr from the seeded stream */
static void fr_mul_plain(u64 r[4], const u64 a[4], const u64 b[4]) {
    /* schoolbook 256x256 then reduce mod r by shift-subtract (setup only) */
    u64 t[8] = {0};
    for (int i = 0; i < 4; i++) {
        u64 c = 0;
        for (int j = 0; j < 4; j++) {
            u128 s = (u128)a[i] * b[j] + t[i + j] + c;
            t[i + j] = (u64)s;
            c = (u64)(s >> 64);
        }
        t[i + 4] = c;
    }
    u64 rem[5] = {0};
    for (int bit = 511; bit >= 0; bit--) {
        for (int k = 4; k > 0; k--) rem[k] = (rem[k] << 1) | (rem[k - 1] >> 63);
        rem[0] = (rem[0] << 1) | ((t[bit / 64] >> (bit % 64)) & 1);
        int ge = rem[4] != 0;
        if (!ge) {
            ge = 1;
            for (int k = 3; k >= 0; k--) {
                if (rem[k] != RORD[k]) { ge = rem[k] > RORD[k]; break; }
            }
        }
        if (ge) {
            u64 borrow = 0;
            for (int k = 0; k < 4; k++) {
                u128 s = (u128)rem[k] - RORD[k] - borrow;
                rem[k] = (u64)s;
                borrow = (u64)(s >> 64) & 1;
            }
            rem[4] -= borrow;
        }
    }
    memcpy(r, rem, 32);
}

void or_gen_srs(u64 *out_affine, size_t n, uint64_t seed, int threads) {
    u64 s = seed ? seed : 1, tau[4], pw[4] = {1, 0, 0, 0};
    draw_fr(&s, tau);
    uint8_t *keys = malloc(n * 32 + 1);
    for (size_t i = 0; i < n; i++) {
        store_le32(keys + 32 * i, pw);
        fr_mul_plain(pw, pw, tau);
    }
    gen_from_keys(out_affine, keys, n, threads);
    free(keys);
}

/* Edge cases at seeded distinct positions: 1% scalar 0, 1% scalar 1,
 * 1% scalar r-1, 0.5% points replaced by copies of other points. */
void or_apply_edge_cases(u64 *points, uint8_t *scalars32, size_t n, uint64_t seed) {
    if (n < 2) return;
    u64 s = seed ? seed : 1;
    size_t *perm = malloc(n * sizeof(size_t));
    for (size_t i = 0; i < n; i++) perm[i] = i;
    size_t n1 = n / 100, n05 = n / 200, total = 3 * n1 + n05;
    if (total > n) total = n;
    for (size_t i = 0; i < total; i++) {
        size_t j = i + (size_t)(rng_next(&s) % (u64)(n - i));
        size_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
    u64 rm1[4];
    memcpy(rm1, RORD, 32);
    rm1[0] -= 1;
    u64 zero[4] = {0, 0, 0, 0}, one[4] = {1, 0, 0, 0};
    size_t k = 0;
    for (size_t i = 0; i < n1; i++) store_le32(scalars32 + 32 * perm[k++], zero);
    for (size_t i = 0; i < n1; i++) store_le32(scalars32 + 32 * perm[k++], one);
    for (size_t i = 0; i < n1; i++) store_le32(scalars32 + 32 * perm[k++], rm1);
    for (size_t i = 0; i < n05 && k < n; i++, k++) {
        size_t dst = perm[k], src = (size_t)(rng_next(&s) % (u64)n);
        if (src == dst) src = (src + 1) % n;
        memcpy(points + 12 * dst, points + 12 * src, 96);
    }
    free(perm);
}

/* ------------------------------------------------- fixed-base wbits tables */

/* points_mult_wbits_precompute_sizeof (multi_scalar.h:244-256) with the
 * c_api.cpp:1310-1316 check: wbits in [2, 14], 96 * n * 2^(wbits-1) bytes,
 * 0 on overflow. */
size_t or_wbits_precompute_sizeof(size_t wbits, size_t npoints) {
    if (wbits < 2 || wbits > 14) return 0;
    if (npoints > SIZE_MAX / 96) return 0;
    const size_t bytes = 96 * npoints;
    if (bytes > (SIZE_MAX >> (wbits - 1))) return 0;
    return bytes << (wbits - 1);
}

/* points_mult_wbits_scratch_sizeof (multi_scalar.h:310-315): one
 * BatchAffineNode<Fp> (affine point + Fp, 144 bytes) per point, at most
 * fixed_table_scratch_capacity<Fp>() = 8192 (multi_scalar.h:218-222). */
size_t or_wbits_scratch_sizeof(size_t npoints) {
    return 144 * (npoints < 8192 ? npoints : 8192);
}

/* point_precompute_row (multi_scalar.cpp:268-286): row[0] = P, row[1] = 2P,
 * then row[i] = row[i-1] + P and row[i+1] = 2 row[j] for i = 2, 4, ...,
 * j = 1, 2, ...; an all-zero P leaves [0, 0, 1] entries, which normalise to
 * the all-zero affine point.  msm_precompute_wbits (multi_scalar.cpp:481-580)
 * then writes the rows to affine, row i at table[i * 2^(wbits-1)]. */
void or_wbits_precompute(u64 *table, size_t wbits, const u64 *points, size_t npoints) {
    const size_t m = (size_t)1 << (wbits - 1);
    jac_t *row = malloc(m * sizeof(jac_t));
    for (size_t i = 0; i < npoints; i++) {
        const aff_t *p = (const aff_t *)(points + 12 * i);
        aff_t *dst = (aff_t *)(table + 12 * m * i);
        if (aff_is_inf(p)) {
            memset(dst, 0, m * sizeof(aff_t));
            continue;
        }
        jac_from_aff(&row[0], p);
        jac_double(&row[1], &row[0]);
        for (size_t k = 2, j = 1; k < m; k += 2, j++) {
            jac_add(&row[k], &row[k - 1], &row[0]);
            jac_double(&row[k + 1], &row[j]);
        }
        or_p1s_to_affine((u64 *)dst, (const u64 *)row, m);
    }
    free(row);
}

/* point_gather_booth_wbits (multi_scalar.h:262-276): magnitude in the low
 * wbits bits of the encoding, sign at bit wbits; 0 gathers infinity. */
static void gather_booth(jac_t *out, const aff_t *row, size_t wbits, u64 enc) {
    const u64 sign = (enc >> wbits) & 1;
    const u64 idx = enc & (((u64)1 << wbits) - 1);
    if (idx == 0) {
        memset(out, 0, sizeof(*out));
        return;
    }
    aff_t a = row[idx - 1];
    if (sign) or_fp_neg(a.y, a.y);
    jac_from_aff(out, &a);
}

/* The Booth encoding of point scalar s in the window of msm_mult_table
 * (multi_scalar.cpp:620-707) whose lowest bit is rem: the top window
 * (rem = nbits - nbits % wbits) holds nbits - rem bits, every lower one
 * wbits, each read with the overlap bit below it except the lowest. */
u64 or_table_booth(const uint8_t *s, size_t nbits, size_t wbits, size_t rem) {
    const size_t cur = nbits - rem < wbits ? nbits - rem : wbits;
    const u64 mask = ((u64)1 << (cur + 1)) - 1;
    const u64 wval = rem ? get_wval_limb(s, rem - 1, cur + 1) & mask : (get_wval_limb(s, 0, cur) << 1) & mask;
    return booth_encode(wval, wbits);
}

/* msm_mult_table: from the top window down, the gathered entries of every
 * point are summed into ret, and ret is doubled wbits times between
 * windows. */
void or_mult_wbits(u64 ret[18], const u64 *table, size_t wbits, size_t npoints, const uint8_t *scalars,
                   size_t nbits) {
    jac_t *r = (jac_t *)ret;
    memset(r, 0, sizeof(*r));
    if (npoints == 0 || nbits == 0 || nbits > 256 || wbits < 2 || wbits > 14) return;
    const size_t m = (size_t)1 << (wbits - 1), nbytes = (nbits + 7) / 8;
    const aff_t *tab = (const aff_t *)table;
    jac_t g;
    for (size_t rem = nbits - nbits % wbits;; rem -= wbits) {
        for (size_t i = 0; i < npoints; i++) {
            gather_booth(&g, tab + m * i, wbits, or_table_booth(scalars + nbytes * i, nbits, wbits, rem));
            jac_add(r, r, &g);
        }
        if (rem == 0) break;
        for (size_t k = 0; k < wbits; k++) jac_double(r, r);
    }
}
