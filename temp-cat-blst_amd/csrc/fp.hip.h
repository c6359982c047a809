// BLS12-381 base field Fp on gfx950: 381-bit Montgomery arithmetic on
// twelve 32-bit limbs, little-endian, R = 2^384.
//
// The limb image is bit-identical to blst's six 64-bit limbs
// (blst/scalar/fields/Fp.h:37-157, R = 2^384 in both), so a blst_fp /
// blst_p1_affine buffer copied to the device is already in this format.
//
// Multiplication is column-wise Montgomery (see fp_mul).  On gfx950
// v_mad_u64_u32 issues at the full VALU rate, so the cost of a product is the
// VALU instruction count: 288 multiply-adds, 288 carry folds and a few dozen
// moves, and nothing else.
// Constants follow blst/scalar/fields/Fp.h:168-192 (p, -1/p mod 2^64, R, R^2),
// restated for 32-bit limbs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bls {

struct fp { uint32_t l[12]; };

// p = 0x1a0111ea...ffffaaab
#define BLS_P0  0xffffaaabu
#define BLS_P1  0xb9feffffu
#define BLS_P2  0xb153ffffu
#define BLS_P3  0x1eabfffeu
#define BLS_P4  0xf6b0f624u
#define BLS_P5  0x6730d2a0u
#define BLS_P6  0xf38512bfu
#define BLS_P7  0x64774b84u
#define BLS_P8  0x434bacd7u
#define BLS_P9  0x4b1ba7b6u
#define BLS_P10 0x397fe69au
#define BLS_P11 0x1a0111eau
// -p^-1 mod 2^32
#define BLS_N0  0xfffcfffdu

__device__ __forceinline__ uint32_t p_limb(int i) {
    constexpr uint32_t P[12] = {BLS_P0, BLS_P1, BLS_P2, BLS_P3, BLS_P4, BLS_P5,
                                BLS_P6, BLS_P7, BLS_P8, BLS_P9, BLS_P10, BLS_P11};
    return P[i];
}

// R mod p (Montgomery one)
__device__ __forceinline__ void fp_one(fp& r) {
    constexpr uint32_t ONE[12] = {0x0002fffd, 0x76090000, 0xc40c0002, 0xebf4000b,
                                  0x53c758ba, 0x5f489857, 0x70525745, 0x77ce5853,
                                  0xa256ec6d, 0x5c071a97, 0xfa80e493, 0x15f65ec3};
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = ONE[i];
}

__device__ __forceinline__ void fp_zero(fp& r) {
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = 0;
}

__device__ __forceinline__ bool fp_is_zero(const fp& a) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) acc |= a.l[i];
    return acc == 0;
}

__device__ __forceinline__ bool fp_eq(const fp& a, const fp& b) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) acc |= a.l[i] ^ b.l[i];
    return acc == 0;
}

// r = (t >= p) ? t - p : t, for t < 2p (twelve limbs, no overflow word)
__device__ __forceinline__ void fp_reduce_once(fp& r, const uint32_t t[12]) {
    uint32_t d[12];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = __builtin_subc(t[i], p_limb(i), borrow, &borrow);
    // borrow == 1  <=>  t < p
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = borrow ? t[i] : d[i];
}

__device__ __forceinline__ void fp_add(fp& r, const fp& a, const fp& b) {
    uint32_t s[12];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
    fp_reduce_once(r, s);   // a + b < 2p < 2^382: no carry out of limb 11
}

__device__ __forceinline__ void fp_sub(fp& r, const fp& a, const fp& b) {
    uint32_t d[12];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = __builtin_subc(a.l[i], b.l[i], borrow, &borrow);
    // if a < b add p back: mask is all-ones on borrow
    const uint32_t mask = 0u - borrow;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = __builtin_addc(d[i], p_limb(i) & mask, c, &c);
}

__device__ __forceinline__ void fp_dbl(fp& r, const fp& a) { fp_add(r, a, a); }

__device__ __forceinline__ void fp_neg(fp& r, const fp& a) {
    // p - a, and 0 stays 0
    uint32_t d[12];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = __builtin_subc(p_limb(i), a.l[i], borrow, &borrow);
    const bool z = fp_is_zero(a);
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = z ? 0u : d[i];
}

// conditional negate: r = neg ? -a : a
__device__ __forceinline__ void fp_cneg(fp& r, const fp& a, bool neg) {
    fp n;
    fp_neg(n, a);
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = neg ? n.l[i] : a.l[i];
}

// acc(64) += x*y with the carry out of bit 64 counted into top.
// One v_mad_u64_u32 writes the 64-bit sum of the product and the aligned
// accumulator pair and its carry-out to VCC; one v_addc folds VCC into top.
__device__ __forceinline__ void mac3(uint64_t& acc, uint32_t& top, uint32_t x, uint32_t y) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(top)
        : "v"(x), "v"(y)
        : "vcc");
}

// Product-scanning (column-wise) Montgomery multiplication with the
// reduction interleaved per column (the "FIPS" ordering of Koc et al.):
// column k gathers a_i*b_(k-i) and m_i*p_(k-i) into a 96-bit accumulator,
// the low word of column k < 12 fixes m_k = lo * (-1/p) mod 2^32, and the
// columns 12..23 are the result words.  288 multiply-adds, each with one
// carry fold, and no per-product register shuffling.
__device__ __forceinline__ void fp_mul(fp& r, const fp& a, const fp& b) {
    uint32_t m[12];
    uint32_t t[12];
    uint64_t acc = 0;
    uint32_t top = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
#pragma unroll
        for (int i = 0; i < k; i++) {
            mac3(acc, top, a.l[i], b.l[k - i]);
            mac3(acc, top, m[i], p_limb(k - i));
        }
        mac3(acc, top, a.l[k], b.l[0]);
        m[k] = (uint32_t)acc * BLS_N0;
        mac3(acc, top, m[k], p_limb(0));
        acc = (acc >> 32) | ((uint64_t)top << 32);
        top = 0;
    }
#pragma unroll
    for (int k = 12; k < 23; k++) {
#pragma unroll
        for (int i = k - 11; i < 12; i++) {
            mac3(acc, top, a.l[i], b.l[k - i]);
            mac3(acc, top, m[i], p_limb(k - i));
        }
        t[k - 12] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)top << 32);
        top = 0;
    }
    t[11] = (uint32_t)acc;   // result < 2p < 2^382: nothing above word 11
    fp_reduce_once(r, t);
}

__device__ __forceinline__ void fp_sqr(fp& r, const fp& a) { fp_mul(r, a, a); }

// a^(p-2) = a^-1 (Fermat), left-to-right over the fixed exponent bits; the
// branch is uniform across the wave.  0 maps to 0.
__device__ __forceinline__ void fp_inv(fp& r, const fp& a) {
    constexpr uint32_t E[12] = {BLS_P0 - 2, BLS_P1, BLS_P2, BLS_P3, BLS_P4, BLS_P5,
                                BLS_P6, BLS_P7, BLS_P8, BLS_P9, BLS_P10, BLS_P11};
    fp acc;
    fp_one(acc);
    const fp base = a;
    for (int bit = 380; bit >= 0; bit--) {
        fp_sqr(acc, acc);
        if ((E[bit >> 5] >> (bit & 31)) & 1u) fp_mul(acc, acc, base);
    }
    r = acc;
}

}  // namespace bls
