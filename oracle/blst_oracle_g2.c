/*
 * TEST INFRASTRUCTURE ONLY -- the oracle's G2 part (see blst_oracle.h).
 *
 * CPU restatement of category-blst 1.0.0's G2 Pippenger MSM: the same
 * templates as G1 (multi_scalar.cpp:730-943, point.cpp) instantiated over
 * Fp2 = Fp[i]/(i^2 + 1) (blst/scalar/fields/Fp2.h), built on the Fp
 * functions of blst_oracle.c.  Layout as blst_types.h:100-112: an Fp2 is
 * (c0, c1), six 64-bit Montgomery limbs each; affine (x, y) is 24 limbs
 * (all-zero = infinity), Jacobian (X, Y, Z) 36 limbs (Z = 0 = infinity).
 */
#include "blst_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

typedef struct { u64 c0[6], c1[6]; } fp2_t;
typedef struct { fp2_t x, y; } aff2_t;
typedef struct { fp2_t x, y, z; } jac2_t;
/* Pointxyzz field order (point.h:347-384): x, y, zzz, zz */
typedef struct { fp2_t x, y, zzz, zz; } xyzz2_t;

/* R mod p (Fp.h:170-175) and p (Fp.h:179-182) */
static const u64 ONE1[6] = {0x760900000002fffdULL, 0xebf4000bc40c0002ULL, 0x5f48985753c758baULL,
                            0x77ce585370525745ULL, 0x5c071a97a256ec6dULL, 0x15f65ec3fa80e493ULL};
static const u64 PMOD[6] = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
                            0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};

/* G2 generator, Montgomery form (point.cpp:114-138) */
static const u64 G2X0[6] = {0xf5f28fa202940a10ULL, 0xb3f5fb2687b4961aULL, 0xa1a893b53e2ae580ULL,
                            0x9894999d1a3caee9ULL, 0x6f67b7631863366bULL, 0x058191924350bcd7ULL};
static const u64 G2X1[6] = {0xa5a9c0759e23f606ULL, 0xaaa0c59dbccd60c3ULL, 0x3bb17e18e2867806ULL,
                            0x1b1ab6cc8541b367ULL, 0xc2b6ed0ef2158547ULL, 0x11922a097360edf3ULL};
static const u64 G2Y0[6] = {0x4c730af860494c4aULL, 0x597cfa1f5e369c5aULL, 0xe7e6856caa0a635aULL,
                            0xbbefb5e96e0d495fULL, 0x07d3a975f0ef25a2ULL, 0x0083fd8e7e80dae5ULL};
static const u64 G2Y1[6] = {0xadc0fc92df64b05dULL, 0x18aa270a2b1461dcULL, 0x86adac6a3be4eba0ULL,
                            0x79495c4ec93da33aULL, 0xe7175850a43ccaedULL, 0x0b2bc2a163de1bf2ULL};

static int fp1_is_zero(const u64 a[6]) {
    u64 t = 0;
    for (int i = 0; i < 6; i++) t |= a[i];
    return t == 0;
}

/* ------------------------------------------------------------------ Fp2 */

static int f2_is_zero(const fp2_t *a) { return fp1_is_zero(a->c0) && fp1_is_zero(a->c1); }
static void f2_add(fp2_t *r, const fp2_t *a, const fp2_t *b) {
    or_fp_add(r->c0, a->c0, b->c0);
    or_fp_add(r->c1, a->c1, b->c1);
}
static void f2_sub(fp2_t *r, const fp2_t *a, const fp2_t *b) {
    or_fp_sub(r->c0, a->c0, b->c0);
    or_fp_sub(r->c1, a->c1, b->c1);
}
static void f2_neg(fp2_t *r, const fp2_t *a) {
    or_fp_neg(r->c0, a->c0);
    or_fp_neg(r->c1, a->c1);
}
/* (a0 + a1 i)(b0 + b1 i) = (a0 b0 - a1 b1) + (a0 b1 + a1 b0) i (Fp2.h mul) */
static void f2_mul(fp2_t *r, const fp2_t *a, const fp2_t *b) {
    u64 t0[6], t1[6], t2[6], t3[6];
    or_fp_mul(t0, a->c0, b->c0);
    or_fp_mul(t1, a->c1, b->c1);
    or_fp_mul(t2, a->c0, b->c1);
    or_fp_mul(t3, a->c1, b->c0);
    or_fp_sub(r->c0, t0, t1);
    or_fp_add(r->c1, t2, t3);
}
static void f2_mul3(fp2_t *r, const fp2_t *a) {
    fp2_t t;
    f2_add(&t, a, a);
    f2_add(r, &t, a);
}
static void f2_mul8(fp2_t *r, const fp2_t *a) {
    fp2_t t;
    f2_add(&t, a, a);
    f2_add(&t, &t, &t);
    f2_add(r, &t, &t);
}
/* 1 / (a0 + a1 i) = (a0 - a1 i) / (a0^2 + a1^2) */
static void f2_inv(fp2_t *r, const fp2_t *a) {
    u64 n[6], t[6], ni[6];
    or_fp_mul(n, a->c0, a->c0);
    or_fp_mul(t, a->c1, a->c1);
    or_fp_add(n, n, t);
    or_fp_inv(ni, n);
    or_fp_mul(r->c0, a->c0, ni);
    or_fp_mul(t, a->c1, ni);
    or_fp_neg(r->c1, t);
}
static void f2_one(fp2_t *r) {
    memcpy(r->c0, ONE1, 48);
    memset(r->c1, 0, 48);
}

void or_fp2_mul(u64 r[12], const u64 a[12], const u64 b[12]) {
    f2_mul((fp2_t *)r, (const fp2_t *)a, (const fp2_t *)b);
}

/* ------------------------------------------------------------ G2 points */

static int aff2_is_inf(const aff2_t *a) { return f2_is_zero(&a->x) && f2_is_zero(&a->y); }
static int xyzz2_is_inf(const xyzz2_t *p) { return f2_is_zero(&p->zz) && f2_is_zero(&p->zzz); }

/* Point::doubling, dbl-2009-l (point.cpp:1898-1925) over Fp2 */
static void jac2_double(jac2_t *p3, const jac2_t *p1) {
    fp2_t A, B, C, t;
    jac2_t o;
    f2_mul(&A, &p1->x, &p1->x);
    f2_mul(&B, &p1->y, &p1->y);
    f2_mul(&C, &B, &B);
    f2_add(&B, &B, &p1->x);
    f2_mul(&B, &B, &B);
    f2_sub(&B, &B, &A);
    f2_sub(&B, &B, &C);
    f2_add(&B, &B, &B);          /* D */
    f2_mul3(&A, &A);             /* E */
    f2_mul(&o.x, &A, &A);
    f2_sub(&o.x, &o.x, &B);
    f2_sub(&o.x, &o.x, &B);      /* X3 = F - 2D */
    f2_add(&t, &p1->z, &p1->z);
    f2_mul(&o.z, &t, &p1->y);    /* Z3 = 2 Y1 Z1 */
    f2_mul8(&C, &C);
    f2_sub(&o.y, &B, &o.x);
    f2_mul(&o.y, &o.y, &A);
    f2_sub(&o.y, &o.y, &C);      /* Y3 = E (D - X3) - 8C */
    *p3 = o;
}

/* Point::dadd (point.cpp:1623-1681) over Fp2, doubling and infinity cases
 * resolved with explicit branches */
static void jac2_add(jac2_t *out, const jac2_t *p1, const jac2_t *p2) {
    if (f2_is_zero(&p1->z)) { *out = *p2; return; }
    if (f2_is_zero(&p2->z)) { *out = *p1; return; }
    fp2_t z1z1, z2z2, u1, u2, s1, s2, t, H, R;
    f2_mul(&z1z1, &p1->z, &p1->z);
    f2_mul(&z2z2, &p2->z, &p2->z);
    f2_mul(&u1, &p1->x, &z2z2);
    f2_mul(&u2, &p2->x, &z1z1);
    f2_mul(&t, &p2->z, &z2z2);
    f2_mul(&s1, &p1->y, &t);
    f2_mul(&t, &p1->z, &z1z1);
    f2_mul(&s2, &p2->y, &t);
    f2_sub(&H, &u2, &u1);
    f2_sub(&R, &s2, &s1);
    if (f2_is_zero(&H)) {
        if (f2_is_zero(&R)) { jac2_double(out, p1); return; }
        memset(out, 0, sizeof(*out));
        return;
    }
    fp2_t HH, HHH, sx, V;
    jac2_t o;
    f2_mul(&o.z, &p1->z, &p2->z);
    f2_mul(&o.z, &o.z, &H);
    f2_mul(&HH, &H, &H);
    f2_mul(&HHH, &HH, &H);
    f2_add(&sx, &u1, &u2);
    f2_mul(&t, &HH, &sx);
    f2_mul(&o.x, &R, &R);
    f2_sub(&o.x, &o.x, &t);
    f2_mul(&V, &HH, &u1);
    f2_sub(&V, &V, &o.x);
    f2_mul(&V, &V, &R);
    f2_mul(&t, &HHH, &s1);
    f2_sub(&o.y, &V, &t);
    *out = o;
}

static void jac2_from_aff(jac2_t *out, const aff2_t *a) {
    out->x = a->x;
    out->y = a->y;
    if (aff2_is_inf(a))
        memset(&out->z, 0, sizeof(out->z));
    else
        f2_one(&out->z);
}

static void jac2_to_aff(aff2_t *out, const jac2_t *p) {
    if (f2_is_zero(&p->z)) { memset(out, 0, sizeof(*out)); return; }
    fp2_t zi, zi2, zi3;
    f2_inv(&zi, &p->z);
    f2_mul(&zi2, &zi, &zi);
    f2_mul(&zi3, &zi2, &zi);
    f2_mul(&out->x, &p->x, &zi2);
    f2_mul(&out->y, &p->y, &zi3);
}

/* Pointxyzz::dadd(xyzz, affine, subtract) (point.cpp:2027-2091) over Fp2 */
static void xyzz2_add_aff(xyzz2_t *p3, const xyzz2_t *p1, const aff2_t *p2, int subtract) {
    if (aff2_is_inf(p2)) { *p3 = *p1; return; }
    if (xyzz2_is_inf(p1)) {
        p3->x = p2->x;
        p3->y = p2->y;
        f2_one(&p3->zz);
        f2_one(&p3->zzz);
        if (subtract) f2_neg(&p3->zzz, &p3->zzz);
        return;
    }
    fp2_t Pv, R;
    f2_mul(&Pv, &p2->x, &p1->zz);
    f2_mul(&R, &p2->y, &p1->zzz);
    if (subtract) f2_neg(&R, &R);
    f2_sub(&Pv, &Pv, &p1->x);
    f2_sub(&R, &R, &p1->y);
    xyzz2_t o;
    if (!f2_is_zero(&Pv)) {
        fp2_t PP, PPP, Q, t;
        f2_mul(&PP, &Pv, &Pv);
        f2_mul(&PPP, &PP, &Pv);
        f2_mul(&Q, &p1->x, &PP);
        f2_mul(&o.x, &R, &R);
        f2_add(&t, &Q, &Q);
        f2_sub(&o.x, &o.x, &PPP);
        f2_sub(&o.x, &o.x, &t);
        f2_sub(&Q, &Q, &o.x);
        f2_mul(&Q, &Q, &R);
        f2_mul(&o.y, &p1->y, &PPP);
        f2_sub(&o.y, &Q, &o.y);
        f2_mul(&o.zz, &p1->zz, &PP);
        f2_mul(&o.zzz, &p1->zzz, &PPP);
    } else if (f2_is_zero(&R)) {
        fp2_t U, S, M;
        f2_add(&U, &p2->y, &p2->y);
        f2_mul(&o.zz, &U, &U);
        f2_mul(&o.zzz, &o.zz, &U);
        f2_mul(&S, &p2->x, &o.zz);
        f2_mul(&M, &p2->x, &p2->x);
        f2_mul3(&M, &M);
        f2_mul(&o.x, &M, &M);
        f2_add(&U, &S, &S);
        f2_sub(&o.x, &o.x, &U);
        f2_mul(&o.y, &o.zzz, &p2->y);
        f2_sub(&S, &S, &o.x);
        f2_mul(&S, &S, &M);
        f2_sub(&o.y, &S, &o.y);
        if (subtract) f2_neg(&o.zzz, &o.zzz);
    } else {
        memset(&o, 0, sizeof(o));
    }
    *p3 = o;
}

/* Pointxyzz::dadd(xyzz, xyzz) (point.cpp:1939-2001) over Fp2 */
static void xyzz2_add(xyzz2_t *p3, const xyzz2_t *p1, const xyzz2_t *p2) {
    if (xyzz2_is_inf(p2)) { *p3 = *p1; return; }
    if (xyzz2_is_inf(p1)) { *p3 = *p2; return; }
    fp2_t U, S, Pv, R;
    f2_mul(&U, &p1->x, &p2->zz);
    f2_mul(&S, &p1->y, &p2->zzz);
    f2_mul(&Pv, &p2->x, &p1->zz);
    f2_mul(&R, &p2->y, &p1->zzz);
    f2_sub(&Pv, &Pv, &U);
    f2_sub(&R, &R, &S);
    xyzz2_t o;
    if (!f2_is_zero(&Pv)) {
        fp2_t PP, PPP, Q, t;
        f2_mul(&PP, &Pv, &Pv);
        f2_mul(&PPP, &PP, &Pv);
        f2_mul(&Q, &U, &PP);
        f2_mul(&o.x, &R, &R);
        f2_add(&t, &Q, &Q);
        f2_sub(&o.x, &o.x, &PPP);
        f2_sub(&o.x, &o.x, &t);
        f2_sub(&Q, &Q, &o.x);
        f2_mul(&Q, &Q, &R);
        f2_mul(&o.y, &S, &PPP);
        f2_sub(&o.y, &Q, &o.y);
        f2_mul(&o.zz, &p1->zz, &p2->zz);
        f2_mul(&o.zzz, &p1->zzz, &p2->zzz);
        f2_mul(&o.zz, &o.zz, &PP);
        f2_mul(&o.zzz, &o.zzz, &PPP);
    } else if (f2_is_zero(&R)) {
        fp2_t V, W, M, t;
        f2_add(&U, &p1->y, &p1->y);
        f2_mul(&V, &U, &U);
        f2_mul(&W, &V, &U);
        f2_mul(&S, &p1->x, &V);
        f2_mul(&M, &p1->x, &p1->x);
        f2_mul3(&M, &M);
        f2_mul(&o.x, &M, &M);
        f2_add(&t, &S, &S);
        f2_sub(&o.x, &o.x, &t);
        f2_mul(&o.y, &W, &p1->y);
        f2_sub(&S, &S, &o.x);
        f2_mul(&S, &S, &M);
        f2_sub(&o.y, &S, &o.y);
        f2_mul(&o.zz, &p1->zz, &V);
        f2_mul(&o.zzz, &p1->zzz, &W);
    } else {
        memset(&o, 0, sizeof(o));
    }
    *p3 = o;
}

/* Pointxyzz::to_Jacobian (point.cpp:2109-2114) */
static void xyzz2_to_jac(jac2_t *out, const xyzz2_t *in) {
    jac2_t o;
    f2_mul(&o.x, &in->x, &in->zz);
    f2_mul(&o.y, &in->y, &in->zzz);
    o.z = in->zz;
    *out = o;
}

void or_p2_generator(u64 out[24]) {
    memcpy(out, G2X0, 48);
    memcpy(out + 6, G2X1, 48);
    memcpy(out + 12, G2Y0, 48);
    memcpy(out + 18, G2Y1, 48);
}
void or_p2_from_affine(u64 out[36], const u64 a[24]) { jac2_from_aff((jac2_t *)out, (const aff2_t *)a); }
void or_p2_add(u64 out[36], const u64 a[36], const u64 b[36]) {
    jac2_add((jac2_t *)out, (const jac2_t *)a, (const jac2_t *)b);
}
void or_p2_double(u64 out[36], const u64 a[36]) { jac2_double((jac2_t *)out, (const jac2_t *)a); }
void or_p2_to_affine(u64 out[24], const u64 p[36]) { jac2_to_aff((aff2_t *)out, (const jac2_t *)p); }

/* variable-time double-and-add over the low nbits (mult_vartime's value) */
void or_p2_mult(u64 out[36], const u64 p[36], const uint8_t *scalar, size_t nbits) {
    jac2_t acc;
    memset(&acc, 0, sizeof(acc));
    for (size_t b = nbits; b-- > 0;) {
        jac2_double(&acc, &acc);
        if ((scalar[b / 8] >> (b % 8)) & 1) jac2_add(&acc, &acc, (const jac2_t *)p);
    }
    memcpy(out, &acc, sizeof(acc));
}

/* points_to_affine over Fp2 (multi_scalar.cpp:160-238): Montgomery's trick */
void or_p2s_to_affine(u64 *out_affine, const u64 *jacobian, size_t n) {
    if (n == 0) return;
    const jac2_t *in = (const jac2_t *)jacobian;
    aff2_t *out = (aff2_t *)out_affine;
    fp2_t *pre = malloc(n * sizeof(fp2_t));
    fp2_t acc, inv;
    f2_one(&acc);
    for (size_t i = 0; i < n; i++) {
        pre[i] = acc;
        if (!f2_is_zero(&in[i].z)) f2_mul(&acc, &acc, &in[i].z);
    }
    f2_inv(&inv, &acc);
    for (size_t i = n; i-- > 0;) {
        if (f2_is_zero(&in[i].z)) {
            memset(&out[i], 0, sizeof(out[i]));
            continue;
        }
        fp2_t zi, zi2, zi3;
        f2_mul(&zi, &inv, &pre[i]);
        f2_mul(&inv, &inv, &in[i].z);
        f2_mul(&zi2, &zi, &zi);
        f2_mul(&zi3, &zi2, &zi);
        f2_mul(&out[i].x, &in[i].x, &zi2);
        f2_mul(&out[i].y, &in[i].y, &zi3);
    }
    free(pre);
}

/* canonical value of a Montgomery Fp element, big-endian */
static void fp_be_bytes(uint8_t out[48], const u64 a[6]) {
    u64 v[6];
    or_fp_from_mont(v, a);
    for (int i = 0; i < 48; i++) out[47 - i] = (uint8_t)(v[i / 8] >> (8 * (i % 8)));
}

static int fp_gt_half(const u64 a[6]) {
    /* (p - 1) / 2 < v, on canonical values */
    u64 v[6], h[6];
    or_fp_from_mont(v, a);
    for (int i = 0; i < 6; i++) h[i] = (PMOD[i] >> 1) | (i < 5 ? PMOD[i + 1] << 63 : 0);
    for (int i = 5; i >= 0; i--)
        if (v[i] != h[i]) return v[i] > h[i];
    return 0;
}

/* compress_BE (point.cpp:874-902) for G2: x.c1 then x.c0, flag 0x80, 0x40
 * for infinity, 0x20 when y is the larger root (its c1, or c0 when c1 = 0) */
void or_p2_compress(uint8_t out[96], const u64 p[36]) {
    const jac2_t *j = (const jac2_t *)p;
    if (f2_is_zero(&j->z)) {
        memset(out, 0, 96);
        out[0] = 0xc0;
        return;
    }
    aff2_t a;
    jac2_to_aff(&a, j);
    fp_be_bytes(out, a.x.c1);
    fp_be_bytes(out + 48, a.x.c0);
    const int big = fp1_is_zero(a.y.c1) ? fp_gt_half(a.y.c0) : fp_gt_half(a.y.c1);
    out[0] |= (uint8_t)(0x80 | (big ? 0x20 : 0));
}

/* ---------------------------------------------------------- Pippenger G2 */

/* point_integrate_buckets (multi_scalar.cpp:730-751) over Fp2 */
static void integrate2(jac2_t *out, xyzz2_t *buckets, size_t count) {
    size_t n = count;
    xyzz2_t acc, ret;
    acc = buckets[--n];
    ret = buckets[n];
    memset(&buckets[n], 0, sizeof(xyzz2_t));
    while (n--) {
        xyzz2_add(&acc, &acc, &buckets[n]);
        xyzz2_add(&ret, &ret, &acc);
        memset(&buckets[n], 0, sizeof(xyzz2_t));
    }
    xyzz2_to_jac(out, &ret);
}

/* points_tile_pippenger (multi_scalar.cpp:806-865) over Fp2; the digit is
 * WindowExtract's (or_booth_digit): magnitude in the low e bits, sign at e */
static void tile2(jac2_t *out, const aff2_t *points, const uint8_t *scalars, size_t n, size_t nbits,
                  xyzz2_t *buckets, size_t bit0, size_t window) {
    const size_t e = nbits - bit0 + 1 < window ? nbits - bit0 + 1 : window;
    const size_t nbytes = (nbits + 7) / 8;
    const u64 mask = ((u64)1 << e) - 1;
    for (size_t i = 0; i < n; i++) {
        const u64 d = or_booth_digit(scalars + i * nbytes, nbits, bit0, window);
        const int sign = (int)((d >> e) & 1);
        const u64 idx = d & mask;
        if (idx) xyzz2_add_aff(&buckets[idx - 1], &buckets[idx - 1], &points[i], sign);
    }
    integrate2(out, buckets, (size_t)1 << (e - 1));
}

void or_msm_p2_tile(u64 ret[36], const u64 *points, const uint8_t *scalars, size_t npoints, size_t nbits,
                    size_t bit0, size_t window) {
    xyzz2_t *buckets = calloc((size_t)1 << (window - 1), sizeof(xyzz2_t));
    tile2((jac2_t *)ret, (const aff2_t *)points, scalars, npoints, nbits, buckets, bit0, window);
    free(buckets);
}

/* msm_mult_pippenger (multi_scalar.cpp:892-925, G2 instantiation :935-943) */
void or_msm_p2_pippenger(u64 ret[36], const u64 *points, const uint8_t *scalars, size_t npoints, size_t nbits,
                         size_t window) {
    jac2_t *r = (jac2_t *)ret;
    memset(r, 0, sizeof(*r));
    if (npoints == 0 || nbits == 0 || nbits > 256) return;
    if (window == 0) window = or_pippenger_window_size(npoints);
    xyzz2_t *buckets = calloc((size_t)1 << (window - 1), sizeof(xyzz2_t));
    const size_t nwin = 1 + nbits / window;
    for (size_t wi = nwin; wi-- > 0;) {
        jac2_t tile;
        tile2(&tile, (const aff2_t *)points, scalars, npoints, nbits, buckets, wi * window, window);
        jac2_add(r, r, &tile);
        if (wi)
            for (size_t k = 0; k < window; k++) jac2_double(r, r);
    }
    free(buckets);
}

/* point_precompute_row over Fp2 (multi_scalar.cpp:268-286): row[0] = P,
 * row[1] = 2P, then row[k] = row[k-1] + P and row[k+1] = 2 row[j] for
 * k = 2, 4, ..., j = 1, 2, ...; an all-zero P gives an all-zero row.
 * msm_precompute_wbits<Fp2> (multi_scalar.cpp:481-580) writes row i, in
 * affine form, at table[i * 2^(wbits-1)]. */
void or_p2_wbits_precompute(u64 *table, size_t wbits, const u64 *points, size_t npoints) {
    const size_t m = (size_t)1 << (wbits - 1);
    jac2_t *row = malloc(m * sizeof(jac2_t));
    for (size_t i = 0; i < npoints; i++) {
        const aff2_t *p = (const aff2_t *)(points + 24 * i);
        aff2_t *dst = (aff2_t *)(table + 24 * m * i);
        if (aff2_is_inf(p)) {
            memset(dst, 0, m * sizeof(aff2_t));
            continue;
        }
        jac2_from_aff(&row[0], p);
        if (m > 1) jac2_double(&row[1], &row[0]);
        for (size_t k = 2, j = 1; k < m; k += 2, j++) {
            jac2_add(&row[k], &row[k - 1], &row[0]);
            jac2_double(&row[k + 1], &row[j]);
        }
        or_p2s_to_affine((u64 *)dst, (const u64 *)row, m);
    }
    free(row);
}

/* point_gather_booth_wbits over Fp2 (multi_scalar.h:262-276) */
static void gather_booth2(jac2_t *out, const aff2_t *row, size_t wbits, u64 enc) {
    const u64 sign = (enc >> wbits) & 1;
    const u64 idx = enc & (((u64)1 << wbits) - 1);
    if (idx == 0) {
        memset(out, 0, sizeof(*out));
        return;
    }
    aff2_t a = row[idx - 1];
    if (sign) f2_neg(&a.y, &a.y);
    jac2_from_aff(out, &a);
}

/* msm_mult_table<Fp2> (multi_scalar.cpp:620-707), as or_mult_wbits */
void or_p2_mult_wbits(u64 ret[36], const u64 *table, size_t wbits, size_t npoints, const uint8_t *scalars,
                      size_t nbits) {
    jac2_t *r = (jac2_t *)ret;
    memset(r, 0, sizeof(*r));
    if (npoints == 0 || nbits == 0 || nbits > 256 || wbits < 2 || wbits > 14) return;
    const size_t m = (size_t)1 << (wbits - 1), nbytes = (nbits + 7) / 8;
    const aff2_t *tab = (const aff2_t *)table;
    jac2_t g;
    for (size_t rem = nbits - nbits % wbits;; rem -= wbits) {
        for (size_t i = 0; i < npoints; i++) {
            gather_booth2(&g, tab + m * i, wbits, or_table_booth(scalars + nbytes * i, nbits, wbits, rem));
            jac2_add(r, r, &g);
        }
        if (rem == 0) break;
        for (size_t k = 0; k < wbits; k++) jac2_double(r, r);
    }
}

/* Synthetic G2 bases: P_i = [a + i b] G2 for a, b drawn from the seeded
 * stream (one addition per point, then one batched normalisation). */
void or_gen_p2_points(u64 *out_affine, size_t n, uint64_t seed) {
    if (n == 0) return;
    uint8_t ab[64];
    or_gen_scalars(ab, 2, seed, 0);
    jac2_t g, A, Bp;
    u64 gen[24];
    or_p2_generator(gen);
    jac2_from_aff(&g, (const aff2_t *)gen);
    or_p2_mult((u64 *)&A, (const u64 *)&g, ab, 255);
    or_p2_mult((u64 *)&Bp, (const u64 *)&g, ab + 32, 255);
    jac2_t *jac = malloc(n * sizeof(jac2_t));
    jac[0] = A;
    for (size_t i = 1; i < n; i++) jac2_add(&jac[i], &jac[i - 1], &Bp);
    or_p2s_to_affine(out_affine, (const u64 *)jac, n);
    free(jac);
}
