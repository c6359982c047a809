// BLS12-381 Fp in the representation of the G1 bucket accumulation:
// fourteen 28-bit limbs (value = sum l_i 2^(28 i)) and Montgomery radix
// 2^392.
//
// Why a second representation: with 32-bit limbs (fp.hip.h) every
// v_mad_u64_u32 of a product can carry out of its 64-bit accumulator, so each
// one is followed by a carry fold, 576 instructions for a product's 288
// multiply-adds.  A 28-bit limb product is below 2^56, and one column of a
// product (at most 14 a_i*b_j and 14 m_i*p_j) stays below 2^64 with room to
// spare, so a product is 392 multiply-adds and no folds.  The spare bits also
// let subtractions skip their reduction ("lazy" values: limbs above 2^28, the
// value above 2p), since the products accept such inputs; the bounds each
// operation needs and gives are stated with it.
//
// Crossing to and from the 32-bit form costs no multiplication:
//   in   a canonical word image w (blst Montgomery form, w = v 2^384 mod p)
//        shifted left by 8 bits is v 2^392: the exact R = 2^392 image of the
//        same element, unreduced (below 2^389; fq_from_words_shl8).
//   out  an fq value c read as a 32-bit-limb word image is v 2^8 (c = v 2^392
//        read with R = 2^384).  An XYZZ point whose X and ZZ are written
//        as is and whose Y and ZZZ are first multiplied by 16 is read as
//        (2^8 X, 2^12 Y, 2^8 ZZ, 2^12 ZZZ): the same point, rescaled by
//        lambda = 2^4 (x = X/ZZ, y = Y/ZZZ and ZZ^3 = ZZZ^2 all hold).
//        The bucket sum keeps ZZZ times 16 throughout (fq_xyzz), which
//        costs nothing: ZZZ only enters S2 = y ZZZ1 = (y/16) (16 ZZZ1), and
//        y/16 is again an exact shifted image (w 2^4), and ZZZ3 = ZZZ1 PPP is
//        linear in ZZZ1.  So only Y is multiplied by 16 on the way out.
// The formulas of the accumulation (madd-2008-s, dbl-2008-s-1; see ec.hip.h)
// are the ones of the 32-bit path.
#pragma once
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BLS_FQ __host__ __device__ __forceinline__
#else
#define BLS_FQ static inline
#endif

namespace bls {

struct fq { uint32_t l[14]; };

constexpr uint32_t kFqMask = 0x0fffffffu;
constexpr uint32_t kFqN0 = 0x0ffcfffdu;        // -p^-1 mod 2^28

// p
BLS_FQ uint32_t fq_p(int i) {
    constexpr uint32_t Q[14] = {0x0fffaaab, 0x0fefffff, 0x03ffffb9, 0x0fffeb15, 0x06241eab, 0x0a0f6b0f, 0x0f6730d2,
                                0x0f38512b, 0x04774b84, 0x04bacd76, 0x0ba7b643, 0x0e69a4b1, 0x01ea397f, 0x0001a011};
    return Q[i];
}

// 4p with every limb below the top one at least 2^28 - 1 (one 2^28 borrowed
// from each next limb): a normalized value below 2p can be subtracted from
// it limb by limb
BLS_FQ uint32_t fq_4p(int i) {
    constexpr uint32_t K[14] = {0x1ffeaaac, 0x1fbffffe, 0x1ffffee6, 0x1fffac53, 0x18907aae, 0x183dac3c, 0x1d9cc349,
                                0x1ce144ae, 0x11dd2e12, 0x12eb35d8, 0x1e9ed90c, 0x19a692c5, 0x17a8e5fe, 0x00068043};
    return K[i];
}

// 6p with every limb below the top one at least 3 (2^28 - 1): a sum of three
// normalized values below 1.1p each can be subtracted from it limb by limb
BLS_FQ uint32_t fq_6p(int i) {
    constexpr uint32_t K[14] = {0x3ffe0002, 0x3f9ffffc, 0x37fffe58, 0x3fff827c, 0x34d8b804, 0x3c5c8259, 0x3c6b24ec,
                                0x3b51e704, 0x3acbc51a, 0x3c60d0c2, 0x35ee4590, 0x3679dc27, 0x3b7d58fc, 0x0009c063};
    return K[i];
}

// 2^392 mod p: one
BLS_FQ void fq_one(fq& r) {
    constexpr uint32_t O[14] = {0x0347fcb8, 0x0d800000, 0x0002b119, 0x00cde6d2, 0x0c7212e0, 0x083a2090, 0x0037669f,
                                0x0da0f73e, 0x09b09b42, 0x01297bb0, 0x0515d98f, 0x0012ca7c, 0x0659fcfa, 0x0000577a};
#pragma unroll
    for (int i = 0; i < 14; i++) r.l[i] = O[i];
}

// 2^-4 2^392 mod p: the element 1/16
BLS_FQ void fq_inv16(fq& r) {
    constexpr uint32_t O[14] = {0x00345521, 0x09d00000, 0x0c002aee, 0x000cd3f7, 0x0bd93084, 0x048b5790, 0x0db70ed3,
                                0x0a763809, 0x02d6af76, 0x096ffe76, 0x0a2538ba, 0x0935ff00, 0x035abc8f, 0x0000d580};
#pragma unroll
    for (int i = 0; i < 14; i++) r.l[i] = O[i];
}

BLS_FQ void fq_zero(fq& r) {
#pragma unroll
    for (int i = 0; i < 14; i++) r.l[i] = 0;
}

// Montgomery product r = a b 2^-392 mod p, product scanning with the
// reduction interleaved per column: column k < 14 fixes m_k from its low 28
// bits, columns 14..26 are the result limbs.
//   in   limbs below 2^29.6 (lazy values included), values A, B
//   out  normalized limbs, value below A B / 2^392 + p (below 1.1p for
//        A, B below 9p; below 1.2p for A below 2^389 and B below 2p)
// Column bound: 14 terms below 2^59.2 and 14 below 2^56, plus the carry in,
// is below 2^63.1.
BLS_FQ void fq_mul(fq& r, const fq& a, const fq& b) {
    uint32_t m[14];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 14; k++) {
#pragma unroll
        for (int i = 0; i < k; i++) {
            acc += (uint64_t)a.l[i] * b.l[k - i];
            acc += (uint64_t)m[i] * fq_p(k - i);
        }
        acc += (uint64_t)a.l[k] * b.l[0];
        m[k] = ((uint32_t)acc * kFqN0) & kFqMask;
        acc += (uint64_t)m[k] * fq_p(0);
        acc >>= 28;
    }
#pragma unroll
    for (int k = 14; k < 27; k++) {
#pragma unroll
        for (int i = k - 13; i < 14; i++) {
            acc += (uint64_t)a.l[i] * b.l[k - i];
            acc += (uint64_t)m[i] * fq_p(k - i);
        }
        r.l[k - 14] = (uint32_t)acc & kFqMask;
        acc >>= 28;
    }
    r.l[13] = (uint32_t)acc;
}

// r = a^2 2^-392: the cross products a_i a_j (i < j) once, against the
// doubled limb 2 a_j, and the squares a_i^2; 105 products where fq_mul has
// 196.  Same bounds as fq_mul (the doubled limbs are below 2^30.6, a column
// holds at most 7 cross products).
BLS_FQ void fq_sqr(fq& r, const fq& a) {
    uint32_t d[14], m[14];
#pragma unroll
    for (int i = 0; i < 14; i++) d[i] = a.l[i] << 1;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 14; k++) {
#pragma unroll
        for (int i = 0; 2 * i < k; i++) acc += (uint64_t)a.l[i] * d[k - i];
        if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
        for (int i = 0; i < k; i++) acc += (uint64_t)m[i] * fq_p(k - i);
        m[k] = ((uint32_t)acc * kFqN0) & kFqMask;
        acc += (uint64_t)m[k] * fq_p(0);
        acc >>= 28;
    }
#pragma unroll
    for (int k = 14; k < 27; k++) {
#pragma unroll
        for (int i = k - 13; 2 * i < k; i++) acc += (uint64_t)a.l[i] * d[k - i];
        if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
        for (int i = k - 13; i < 14; i++) acc += (uint64_t)m[i] * fq_p(k - i);
        r.l[k - 14] = (uint32_t)acc & kFqMask;
        acc >>= 28;
    }
    r.l[13] = (uint32_t)acc;
}

// r = (a b + c d) 2^-392 with one reduction.
//   in   a, b limbs below 2^29.6; c limbs below 2^29.1, d normalized
//   out  normalized, value below (A B + C D) / 2^392 + p
// Column bound: 14 terms below 2^59.2, 14 below 2^57.1 and 14 below 2^56:
// below 2^63.5.
BLS_FQ void fq_mul_sum(fq& r, const fq& a, const fq& b, const fq& c, const fq& d) {
    uint32_t m[14];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 14; k++) {
#pragma unroll
        for (int i = 0; i < k; i++) {
            acc += (uint64_t)a.l[i] * b.l[k - i];
            acc += (uint64_t)c.l[i] * d.l[k - i];
            acc += (uint64_t)m[i] * fq_p(k - i);
        }
        acc += (uint64_t)a.l[k] * b.l[0];
        acc += (uint64_t)c.l[k] * d.l[0];
        m[k] = ((uint32_t)acc * kFqN0) & kFqMask;
        acc += (uint64_t)m[k] * fq_p(0);
        acc >>= 28;
    }
#pragma unroll
    for (int k = 14; k < 27; k++) {
#pragma unroll
        for (int i = k - 13; i < 14; i++) {
            acc += (uint64_t)a.l[i] * b.l[k - i];
            acc += (uint64_t)c.l[i] * d.l[k - i];
            acc += (uint64_t)m[i] * fq_p(k - i);
        }
        r.l[k - 14] = (uint32_t)acc & kFqMask;
        acc >>= 28;
    }
    r.l[13] = (uint32_t)acc;
}

// r = a - b + 4p, limb by limb (b normalized, value below 2p): limbs below
// 2^28 + 2^29 for a normalized a, value below A + 4p
BLS_FQ void fq_sub_lazy(fq& r, const fq& a, const fq& b) {
#pragma unroll
    for (int i = 0; i < 14; i++) r.l[i] = a.l[i] + fq_4p(i) - b.l[i];
}

// r = 4p - a (a normalized, below 2p): limbs below 2^29.1, value below 4p
BLS_FQ void fq_neg_lazy(fq& r, const fq& a) {
#pragma unroll
    for (int i = 0; i < 14; i++) r.l[i] = fq_4p(i) - a.l[i];
}

// r = x mod p up to one p: x with limbs below 2^30.4 and value V below 8p
// gives normalized r = V - q p in [0, 2p).  q comes from the top limbs
// (x_13 + floor(x_12 / 2^28)) / (p_13 + 1), which is floor(V / p) or one less
// (the limbs below x_12 add less than 2^339 to V), and at most 7, so every
// x_i - q p_i + carry stays in int32.
BLS_FQ void fq_reduce(fq& r, const fq& x) {
    const uint32_t top = x.l[13] + (x.l[12] >> 28);
    const uint32_t q = top / (fq_p(13) + 1);
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 14; i++) {
        const int32_t t = (int32_t)(x.l[i] - q * fq_p(i)) + c;
        r.l[i] = (uint32_t)t & kFqMask;
        c = t >> 28;
    }
}

// r = x mod p up to one p for a normalized x of any value: q =
// floor(x_13 / (p_13 + 1)) is below 2^12 and q p <= q (p_13 + 1) 2^364 <= V,
// while V < (q + 1)(p_13 + 1) 2^364 and q p >= q p_13 2^364 leave
// V - q p < (p_13 + 1 + q) 2^364 < 1.03 p.  So r is normalized, below 1.03p
// (below 1.01p for V below 2^390, where q < 631).  q p_i < 2^40 needs 64-bit
// differences.
BLS_FQ void fq_reduce_wide(fq& r, const fq& x) {
    const uint32_t q = x.l[13] / (fq_p(13) + 1);
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 14; i++) {
        const int64_t t = (int64_t)x.l[i] - (int64_t)((uint64_t)q * fq_p(i)) + c;
        r.l[i] = (uint32_t)t & kFqMask;
        c = t >> 28;
    }
}

// the canonical value in [0, p) of a normalized value below 2p
BLS_FQ void fq_canon(fq& r, const fq& a) {
    fq t;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 14; i++) {
        const int32_t v = (int32_t)a.l[i] - (int32_t)fq_p(i) + c;
        t.l[i] = (uint32_t)v & kFqMask;
        c = v >> 28;
    }
    const bool ge = c >= 0;
#pragma unroll
    for (int i = 0; i < 14; i++) r.l[i] = ge ? t.l[i] : a.l[i];
}

// X3 = R2 - (PPP + 2Q) (or R2 - 2Q for a doubling) reduced to [0, 2p):
// R2, PPP, Q normalized and below 1.1p; R2 + 6p - S has limbs below 2^30.4
// and value below 7.1p
BLS_FQ void fq_x3(fq& r, const fq& R2, const fq& PPP, const fq& Q, bool dbl) {
    fq t;
#pragma unroll
    for (int i = 0; i < 14; i++) {
        const uint32_t s = (dbl ? 0u : PPP.l[i]) + (Q.l[i] << 1);
        t.l[i] = R2.l[i] + fq_6p(i) - s;
    }
    fq_reduce(r, t);
}

// r = 4 a mod p up to one p (a normalized below 2p)
BLS_FQ void fq_times4(fq& r, const fq& a) {
    fq t;
#pragma unroll
    for (int i = 0; i < 14; i++) t.l[i] = a.l[i] << 2;
    fq_reduce(r, t);
}

// zero modulo p for a normalized value below 2p: 0 or p
BLS_FQ bool fq_is_zero(const fq& a) {
    uint32_t z = 0, q = 0;
#pragma unroll
    for (int i = 0; i < 14; i++) {
        z |= a.l[i];
        q |= a.l[i] ^ fq_p(i);
    }
    return z == 0 || q == 0;
}

// (hi:lo) >> s for s in [0, 32)
BLS_FQ uint32_t fq_funnel(uint32_t hi, uint32_t lo, int s) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)s);
#else
    return s == 0 ? lo : (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#endif
}

// limbs of w 2^sh (w: twelve 32-bit words, value below 2^384 / 2^sh)
template <int SH>
BLS_FQ void fq_from_words(fq& r, const uint32_t w[12]) {
#pragma unroll
    for (int i = 0; i < 14; i++) {
        const int lo = 28 * i - SH;
        if (lo < 0) {
            r.l[i] = (w[0] << -lo) & kFqMask;
        } else {
            const int j = lo >> 5, s = lo & 31;
            const uint32_t hi = j + 1 < 12 ? w[j + 1] : 0u;
            r.l[i] = fq_funnel(hi, w[j], s) & kFqMask;
        }
    }
}

// twelve 32-bit words of a normalized value below 2^384
BLS_FQ void fq_to_words(uint32_t w[12], const fq& a) {
#pragma unroll
    for (int j = 0; j < 12; j++) {
        const int bit = 32 * j, i = bit / 28, s = bit % 28;
        uint32_t v = a.l[i] >> s;
        v |= a.l[i + 1] << (28 - s);
        if (56 - s < 32) v |= a.l[i + 2] << (56 - s);
        w[j] = v;
    }
}

// 16 (2^4 2^392 mod p): ZZZ of an infinite bucket sum
BLS_FQ void fq_sixteen(fq& r) {
    constexpr uint32_t O[14] = {0x0480cb7f, 0x08300003, 0x042b126f, 0x0cdeabe0, 0x04b4d1fc, 0x0573c7de, 0x0540d780,
                                0x0c66805c, 0x0da3d19e, 0x046752a6, 0x0e667627, 0x05efb9af, 0x0fe12320, 0x00009772};
#pragma unroll
    for (int i = 0; i < 14; i++) r.l[i] = O[i];
}

// XYZZ bucket sum in fq: (X, Y, ZZ, 16 ZZZ).  inf is kept as a flag, and an
// infinite sum holds ZZ = 1 and 16 ZZZ = 16 so that the next addition's first
// two products give the incoming point's coordinates reduced (U2 = x,
// S2 = (y/16) 16 = y).
struct fq_xyzz {
    fq x, y, zz, zzz16;
    bool inf;
};

BLS_FQ void fq_xyzz_set_inf(fq_xyzz& p) {
    p.inf = true;
    fq_zero(p.x);
    fq_zero(p.y);
    fq_one(p.zz);
    fq_sixteen(p.zzz16);
}

// acc += (x, y/16): x the exact image of an affine x (fq_from_words<8> of
// words below 2p: normalized, below 2^390), y16 the exact image of y/16
// (fq_from_words<4>: below 2^386); the point is not infinity.
// madd-2008-s as in ec.hip.h.  P = 0 (the same x) is found after the fact by
// ZZ3 = ZZ1 P^2 = 0; then PP, PPP, ZZ3 and Q are computed again for the
// doubling (P' = 2 Y1, R' = 3 X1^2, X3 = R'^2 - 2Q') when R = 0 too, and the
// sum becomes infinity otherwise.  The code is straight-line apart from that
// rare branch.
BLS_FQ void fq_xyzz_madd(fq_xyzz& acc, const fq& x, const fq& y16) {
    fq P, R;
    fq_mul(P, x, acc.zz);       // U2
    fq_mul(R, y16, acc.zzz16);  // S2
    if (acc.inf) {
        acc.x = P;
        acc.y = R;
        acc.inf = false;
        return;
    }
    fq_sub_lazy(P, P, acc.x);   // P = U2 - X1
    fq_sub_lazy(R, R, acc.y);   // R = S2 - Y1
    fq PP, PPP, ZZ3, Q, t;
    fq_sqr(PP, P);
    fq_mul(PPP, PP, P);
    fq_mul(ZZ3, acc.zz, PP);
    fq_mul(Q, acc.x, PP);
    bool dbl = false, vanish = false;
    if (fq_is_zero(ZZ3)) {
        fq_reduce(t, R);
        dbl = fq_is_zero(t);
#pragma unroll
        for (int i = 0; i < 14; i++) P.l[i] = acc.y.l[i] << 1;
        fq_sqr(t, acc.x);
#pragma unroll
        for (int i = 0; i < 14; i++) R.l[i] = 3 * t.l[i];
        fq_sqr(PP, P);
        fq_mul(PPP, PP, P);
        fq_mul(ZZ3, acc.zz, PP);
        fq_mul(Q, acc.x, PP);
        vanish = !dbl || fq_is_zero(ZZ3);
    }
    acc.zz = ZZ3;
    fq_mul(acc.zzz16, acc.zzz16, PPP);
    fq_sqr(t, R);
    fq_x3(acc.x, t, PPP, Q, dbl);
    fq_sub_lazy(Q, Q, acc.x);
    fq_neg_lazy(t, acc.y);
    fq_mul_sum(acc.y, Q, R, t, PPP);
    if (vanish) fq_xyzz_set_inf(acc);
}

// the bucket sum as the word image of the 32-bit path's XYZZ point (values
// below 2p; infinity has ZZ = 0), rescaled by lambda = 2^4 as described
// above.  Branch-free, so that the word arrays stay in registers.
BLS_FQ void fq_xyzz_to_words(uint32_t x[12], uint32_t y[12], uint32_t zz[12], uint32_t zzz[12],
                             const fq_xyzz& p) {
    fq t;
    fq_to_words(x, p.x);
    fq_to_words(zz, p.zz);
    fq_to_words(zzz, p.zzz16);
    fq_times4(t, p.y);
    fq_times4(t, t);
    fq_to_words(y, t);
    const uint32_t keep = p.inf ? 0u : ~0u;
#pragma unroll
    for (int i = 0; i < 12; i++) zz[i] &= keep;
}

// the inverse: a 32-bit path XYZZ word image (values below 2p) read with
// R = 2^392 is the point's coordinates times 2^-8 (for ZZZ that is 16 ZZZ
// rescaled by lambda = 2^-4), and Y times 2^-4 more completes the rescaling
BLS_FQ void fq_xyzz_from_words(fq_xyzz& p, const uint32_t x[12], const uint32_t y[12], const uint32_t zz[12],
                               const uint32_t zzz[12]) {
    fq c, t;
    fq_inv16(c);
    fq_from_words<0>(p.zz, zz);
    fq_from_words<0>(p.x, x);
    fq_from_words<0>(p.zzz16, zzz);
    fq_from_words<0>(t, y);
    fq_mul(p.y, t, c);
    p.inf = false;
    if (fq_is_zero(p.zz)) fq_xyzz_set_inf(p);
}

}  // namespace bls
