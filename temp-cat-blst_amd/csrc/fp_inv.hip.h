// Constant-time inversion modulo p by Bernstein-Yang divsteps (D. J.
// Bernstein, B.-Y. Yang, "Fast constant-time gcd computation and modular
// inversion", TCHES 2019(3)), in batches of 30 divsteps on 32-bit words with
// the full-size values in signed 30-bit limbs.  The replacement for the
// Fermat power a^(p-2) (blst's ct_inverse_mod_383 plays this role in
// blst/scalar/fields/Fp.h): ~1/8 of its VALU instructions on gfx950.
//
// divstep(delta, f, g), f odd:
//   delta > 0 and g odd:  (1 - delta, g, (g - f) / 2)
//   g odd:                (1 + delta, f, (g + f) / 2)
//   g even:               (1 + delta, f, g / 2)
// Starting from (1, p, x), Theorem 11.2 of the paper bounds the number of
// divsteps after which g = 0 by floor((49 d + 57) / 17) for
// f^2 + 4 g^2 <= 5 * 2^(2d), d >= 46: 1101 for d = 381, so 37 batches of 30
// (1110) always finish, and then f = +-gcd(p, x) = +-1 for x != 0.
//
// Alongside f and g the batches carry d and e with f = d x, g = e x (mod p),
// from d = 0, e = 1, so the inverse is sign(f) * d.  A batch of 30 divsteps
// acts on the low words only; its matrix M = (u v; q r), with every row's
// |u| + |v| <= 2^30, gives 2^30 (f', g') = M (f, g) exactly, and
// d' = (u d + v e) / 2^30 mod p is made exact by adding a multiple of p.
//
// The functions are host and device code (the host build is only for
// checking them on the CPU).
#pragma once
#include <stdint.h>

#ifndef BLS_HD
#define BLS_HD __host__ __device__ __forceinline__
#endif

namespace bls {
namespace divstep {

constexpr int kLimbs = 13;                  // 13 x 30 = 390 bits, top limb signed
constexpr uint32_t kMask = (1u << 30) - 1;
constexpr int kBatches = 37;                // 37 x 30 = 1110 >= 1101 divsteps
constexpr uint32_t kPinv = 0x30003u;        // p^-1 mod 2^30

// a signed integer sum_i v[i] 2^(30 i), v[0..11] in [0, 2^30), v[12] signed
struct s30 {
    int32_t v[kLimbs];
};

BLS_HD int32_t p30(int i) {
    constexpr int32_t P[kLimbs] = {0x3fffaaab, 0x27fbffff, 0x153ffffb, 0x2affffac, 0x30f6241e,
                                   0x034a83da, 0x112bf673, 0x12e13ce1, 0x2cd76477, 0x1ed90d2e,
                                   0x29a4b1ba, 0x3a8e5ff9, 0x001a0111};
    return P[i];
}

// 30 divsteps on f mod 2^32 and g mod 2^32: step i needs bit 0 of g_i, which
// the low 32 - i bits of the words still give.  Returns the new delta and
// the batch matrix; the matrix entries are formed mod 2^32 (|entry| <= 2^30).
BLS_HD int32_t batch(int32_t delta, uint32_t f, uint32_t g, int32_t& u_out, int32_t& v_out,
                     int32_t& q_out, int32_t& r_out) {
    // rows: 2^i f_i = u f_0 + v g_0, 2^i g_i = q f_0 + r g_0
    uint32_t u = 1, v = 0, q = 0, r = 1;
#pragma unroll
    for (int i = 0; i < 30; i++) {
        const uint32_t odd = 0u - (g & 1u);                                 // all ones if g odd
        const uint32_t swap = odd & (uint32_t)((int32_t)(0 - delta) >> 31);  // and delta > 0
        // g += (swap ? -f : f) when g is odd, with the row (q, r) alike
        g += ((f ^ swap) - swap) & odd;
        q += ((u ^ swap) - swap) & odd;
        r += ((v ^ swap) - swap) & odd;
        // on a swap f takes the old g = g - (-f), and (u, v) the old (q, r)
        f += g & swap;
        u += q & swap;
        v += r & swap;
        delta = (int32_t)(((uint32_t)delta ^ swap) - swap) + 1;            // 1 - delta or 1 + delta
        g >>= 1;                                                            // g is even here
        u <<= 1;
        v <<= 1;
    }
    u_out = (int32_t)u;
    v_out = (int32_t)v;
    q_out = (int32_t)q;
    r_out = (int32_t)r;
    return delta;
}

// (f, g) <- (u f + v g, q f + r g) / 2^30, exact
BLS_HD void update_fg(s30& f, s30& g, int32_t u, int32_t v, int32_t q, int32_t r) {
    int64_t cf = (int64_t)u * f.v[0] + (int64_t)v * g.v[0];
    int64_t cg = (int64_t)q * f.v[0] + (int64_t)r * g.v[0];
    cf >>= 30;   // the low 30 bits are zero
    cg >>= 30;
#pragma unroll
    for (int i = 1; i < kLimbs; i++) {
        cf += (int64_t)u * f.v[i] + (int64_t)v * g.v[i];
        cg += (int64_t)q * f.v[i] + (int64_t)r * g.v[i];
        f.v[i - 1] = (int32_t)((uint32_t)cf & kMask);
        g.v[i - 1] = (int32_t)((uint32_t)cg & kMask);
        cf >>= 30;
        cg >>= 30;
    }
    f.v[kLimbs - 1] = (int32_t)cf;
    g.v[kLimbs - 1] = (int32_t)cg;
}

// (d, e) <- (u d + v e, q d + r e) / 2^30 mod p, for d, e in (-2p, p).
// A negative d or e is first lifted by p (into (-p, p)); the multiple of p
// that clears the low 30 bits is then taken in (-2^30, 0], so
// 2^30 d' lies in (-2^30 p - 2^30 p, 2^30 p) and d' stays in (-2p, p).
BLS_HD void update_de(s30& d, s30& e, int32_t u, int32_t v, int32_t q, int32_t r) {
    const int32_t sd = d.v[kLimbs - 1] >> 31;   // all ones if d < 0
    const int32_t se = e.v[kLimbs - 1] >> 31;
    int32_t md = (u & sd) + (v & se);
    int32_t me = (q & sd) + (r & se);
    int64_t cd = (int64_t)u * d.v[0] + (int64_t)v * e.v[0];
    int64_t ce = (int64_t)q * d.v[0] + (int64_t)r * e.v[0];
    // cd + md p = 0 mod 2^30 needs md = -(cd p^-1) mod 2^30 on top of the lift
    md -= (int32_t)((kPinv * (uint32_t)cd + (uint32_t)md) & kMask);
    me -= (int32_t)((kPinv * (uint32_t)ce + (uint32_t)me) & kMask);
    cd += (int64_t)p30(0) * md;
    ce += (int64_t)p30(0) * me;
    cd >>= 30;
    ce >>= 30;
#pragma unroll
    for (int i = 1; i < kLimbs; i++) {
        cd += (int64_t)u * d.v[i] + (int64_t)v * e.v[i] + (int64_t)p30(i) * md;
        ce += (int64_t)q * d.v[i] + (int64_t)r * e.v[i] + (int64_t)p30(i) * me;
        d.v[i - 1] = (int32_t)((uint32_t)cd & kMask);
        e.v[i - 1] = (int32_t)((uint32_t)ce & kMask);
        cd >>= 30;
        ce >>= 30;
    }
    d.v[kLimbs - 1] = (int32_t)cd;
    e.v[kLimbs - 1] = (int32_t)ce;
}

// a += p when a < 0 (carries resolved)
BLS_HD void add_p_if_negative(s30& a) {
    const int32_t neg = a.v[kLimbs - 1] >> 31;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < kLimbs - 1; i++) {
        const int32_t t = a.v[i] + (p30(i) & neg) + c;
        a.v[i] = (int32_t)((uint32_t)t & kMask);
        c = t >> 30;
    }
    a.v[kLimbs - 1] += (p30(kLimbs - 1) & neg) + c;
}

// a <- -a when neg is all ones (carries resolved)
BLS_HD void cond_negate(s30& a, int32_t neg) {
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < kLimbs - 1; i++) {
        const int32_t t = ((a.v[i] ^ neg) - neg) + c;
        a.v[i] = (int32_t)((uint32_t)t & kMask);
        c = t >> 30;
    }
    a.v[kLimbs - 1] = ((a.v[kLimbs - 1] ^ neg) - neg) + c;
}

// x^-1 mod p for 0 < x < p, as a value in [0, 2p), and 0 for x = 0 (plain
// integers, not Montgomery form; x must be canonical).  Twelve little-endian
// 32-bit words in and out.
BLS_HD void inverse(uint32_t out[12], const uint32_t x[12]) {
    s30 f, g, d, e;
#pragma unroll
    for (int i = 0; i < kLimbs; i++) {
        const int bit = 30 * i, w = bit >> 5, sh = bit & 31;
        uint64_t win = x[w];
        if (w + 1 < 12) win |= (uint64_t)x[w + 1] << 32;
        g.v[i] = (int32_t)((uint32_t)(win >> sh) & kMask);
        f.v[i] = p30(i);
        d.v[i] = 0;
        e.v[i] = 0;
    }
    e.v[0] = 1;
    int32_t delta = 1;
#pragma unroll 1
    for (int b = 0; b < kBatches; b++) {
        int32_t u, v, q, r;
        const uint32_t f0 = (uint32_t)f.v[0] | ((uint32_t)f.v[1] << 30);
        const uint32_t g0 = (uint32_t)g.v[0] | ((uint32_t)g.v[1] << 30);
        delta = batch(delta, f0, g0, u, v, q, r);
        update_fg(f, g, u, v, q, r);
        update_de(d, e, u, v, q, r);
    }
    // g = 0 and f = +-1 (x != 0) or f = p (x = 0, where d = 0): the inverse is
    // sign(f) d, in (-2p, 2p); two conditional additions of p give [0, 2p)
    cond_negate(d, f.v[kLimbs - 1] >> 31);
    add_p_if_negative(d);
    add_p_if_negative(d);
    uint64_t acc = 0;
    int bits = 0, w = 0;
#pragma unroll
    for (int i = 0; i < kLimbs; i++) {
        acc |= (uint64_t)(uint32_t)d.v[i] << bits;
        bits += 30;
        if (bits >= 32 && w < 12) {
            out[w++] = (uint32_t)acc;
            acc >>= 32;
            bits -= 32;
        }
    }
    if (w < 12) out[w] = (uint32_t)acc;
}

}  // namespace divstep
}  // namespace bls
