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
#include "fp_inv.hip.h"

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

__device__ __forceinline__ uint32_t p2_limb(int i) {
    constexpr uint32_t P2[12] = {0xffff5556u, 0x73fdffffu, 0x62a7ffffu, 0x3d57fffdu, 0xed61ec48u, 0xce61a541u,
                                 0xe70a257eu, 0xc8ee9709u, 0x869759aeu, 0x96374f6cu, 0x72ffcd34u, 0x340223d4u};
    return P2[i];
}

// Values live in [0, 2p) ("2p-redundant"): fp_mul leaves out the final
// subtraction (inputs below 2p keep the Montgomery output below 2p because
// 4p < 2^384), additions and subtractions reduce modulo 2p, and a value is
// brought to [0, p) only where it leaves the device (fp_canon).

// r = (t >= p) ? t - p : t, for t < 2p: the canonical representative
__device__ __forceinline__ void fp_canon(fp& r, const fp& t) {
    uint32_t d[12];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = __builtin_subc(t.l[i], p_limb(i), borrow, &borrow);
    // borrow == 1  <=>  t < p
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = borrow ? t.l[i] : d[i];
}

// zero modulo p: 0 or p
__device__ __forceinline__ bool fp_is_zero(const fp& a) {
    uint32_t z = 0, q = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) {
        z |= a.l[i];
        q |= a.l[i] ^ p_limb(i);
    }
    return z == 0 || q == 0;
}

// r = (t >= 2p) ? t - 2p : t, for t < 4p
__device__ __forceinline__ void fp_reduce_2p(fp& r, const uint32_t t[12]) {
    uint32_t d[12];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = __builtin_subc(t[i], p2_limb(i), borrow, &borrow);
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = borrow ? t[i] : d[i];
}

__device__ __forceinline__ void fp_add(fp& r, const fp& a, const fp& b) {
    uint32_t s[12];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
    fp_reduce_2p(r, s);   // a + b < 4p < 2^383: no carry out of limb 11
}

__device__ __forceinline__ void fp_sub(fp& r, const fp& a, const fp& b) {
    uint32_t d[12];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = __builtin_subc(a.l[i], b.l[i], borrow, &borrow);
    // if a < b add 2p back: mask is all-ones on borrow
    const uint32_t mask = 0u - borrow;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = __builtin_addc(d[i], p2_limb(i) & mask, c, &c);
}

__device__ __forceinline__ void fp_dbl(fp& r, const fp& a) { fp_add(r, a, a); }

__device__ __forceinline__ void fp_neg(fp& r, const fp& a) {
    // 2p - a, and 0 stays 0
    uint32_t d[12];
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = __builtin_subc(p2_limb(i), a.l[i], borrow, &borrow);
    uint32_t z = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) z |= a.l[i];
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = z ? d[i] : 0u;
}

// conditional negate: r = neg ? -a : a
__device__ __forceinline__ void fp_cneg(fp& r, const fp& a, bool neg) {
    fp n;
    fp_neg(n, a);
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = neg ? n.l[i] : a.l[i];
}

// Multiply-accumulate groups: acc(64) += sum of x_j*y_j with every carry out
// of bit 64 counted into top.  One v_mad_u64_u32 adds a product to the
// aligned accumulator pair and writes its carry to VCC, one v_addc folds VCC
// into top.  Several products share one asm statement so the compiler's
// hazard padding between statements is paid once per group; the p limbs go
// in scalar registers.  The _new forms start the column's top at the first
// carry instead of adding to it.
__device__ __forceinline__ void mac1(uint64_t& acc, uint32_t& top, uint32_t x, uint32_t y) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(top)
        : "v"(x), "v"(y)
        : "vcc");
}

__device__ __forceinline__ void mac1_ps(uint64_t& acc, uint32_t& top, uint32_t x, uint32_t py) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(top)
        : "v"(x), "s"(py)
        : "vcc");
}

// acc += a0*b0 + m0*p0 + a1*b1 + m1*p1 (p0, p1 constants)
__device__ __forceinline__ void mac4(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t m0,
                                     uint32_t p0, uint32_t a1, uint32_t b1, uint32_t m1, uint32_t p1) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %4, %5, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %6, %7, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %8, %9, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(top)
        : "v"(a0), "v"(b0), "v"(m0), "s"(p0), "v"(a1), "v"(b1), "v"(m1), "s"(p1)
        : "vcc");
}

__device__ __forceinline__ void mac4_new(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t m0,
                                         uint32_t p0, uint32_t a1, uint32_t b1, uint32_t m1, uint32_t p1) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_cndmask_b32_e64 %1, 0, 1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %4, %5, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %6, %7, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %8, %9, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "=&v"(top)
        : "v"(a0), "v"(b0), "v"(m0), "s"(p0), "v"(a1), "v"(b1), "v"(m1), "s"(p1)
        : "vcc");
}

#define BLS_MAC8(FIRST)                                         \
    "v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t" FIRST               \
    "v_mad_u64_u32 %0, vcc, %4, %5, %0\n\t"                     \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %6, %7, %0\n\t"                     \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %8, %9, %0\n\t"                     \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %10, %11, %0\n\t"                   \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %12, %13, %0\n\t"                   \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %14, %15, %0\n\t"                   \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %16, %17, %0\n\t"                   \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"

// four pairs a_j*b_j + m_j*p_j
__device__ __forceinline__ void mac8(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t m0,
                                     uint32_t p0, uint32_t a1, uint32_t b1, uint32_t m1, uint32_t p1,
                                     uint32_t a2, uint32_t b2, uint32_t m2, uint32_t p2, uint32_t a3,
                                     uint32_t b3, uint32_t m3, uint32_t p3) {
    asm(BLS_MAC8("v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t")
        : "+v"(acc), "+v"(top)
        : "v"(a0), "v"(b0), "v"(m0), "s"(p0), "v"(a1), "v"(b1), "v"(m1), "s"(p1),
          "v"(a2), "v"(b2), "v"(m2), "s"(p2), "v"(a3), "v"(b3), "v"(m3), "s"(p3)
        : "vcc");
}

__device__ __forceinline__ void mac8_new(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t m0,
                                         uint32_t p0, uint32_t a1, uint32_t b1, uint32_t m1, uint32_t p1,
                                         uint32_t a2, uint32_t b2, uint32_t m2, uint32_t p2, uint32_t a3,
                                         uint32_t b3, uint32_t m3, uint32_t p3) {
    asm(BLS_MAC8("v_cndmask_b32_e64 %1, 0, 1, vcc\n\t")
        : "+v"(acc), "=&v"(top)
        : "v"(a0), "v"(b0), "v"(m0), "s"(p0), "v"(a1), "v"(b1), "v"(m1), "s"(p1),
          "v"(a2), "v"(b2), "v"(m2), "s"(p2), "v"(a3), "v"(b3), "v"(m3), "s"(p3)
        : "vcc");
}

#undef BLS_MAC8

// acc += a0*b0 + m0*p0
__device__ __forceinline__ void mac2(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t m0,
                                     uint32_t p0) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %4, %5, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(top)
        : "v"(a0), "v"(b0), "v"(m0), "s"(p0)
        : "vcc");
}

__device__ __forceinline__ void mac2_new(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t m0,
                                         uint32_t p0) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_cndmask_b32_e64 %1, 0, 1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %4, %5, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "=&v"(top)
        : "v"(a0), "v"(b0), "v"(m0), "s"(p0)
        : "vcc");
}

// the column's pairs a_i*b_(k-i) + m_i*p_(k-i) for i in [lo, hi)
template <int K, int LO, int HI>
__device__ __forceinline__ void mac_pairs(uint64_t& acc, uint32_t& top, const fp& a, const fp& b,
                                          const uint32_t m[12]) {
    constexpr bool opens = LO == 0 || (K >= 12 && LO == K - 11);
    if constexpr (HI - LO >= 4) {
        if constexpr (opens)
            mac8_new(acc, top, a.l[LO], b.l[K - LO], m[LO], p_limb(K - LO), a.l[LO + 1], b.l[K - LO - 1],
                     m[LO + 1], p_limb(K - LO - 1), a.l[LO + 2], b.l[K - LO - 2], m[LO + 2], p_limb(K - LO - 2),
                     a.l[LO + 3], b.l[K - LO - 3], m[LO + 3], p_limb(K - LO - 3));
        else
            mac8(acc, top, a.l[LO], b.l[K - LO], m[LO], p_limb(K - LO), a.l[LO + 1], b.l[K - LO - 1],
                 m[LO + 1], p_limb(K - LO - 1), a.l[LO + 2], b.l[K - LO - 2], m[LO + 2], p_limb(K - LO - 2),
                 a.l[LO + 3], b.l[K - LO - 3], m[LO + 3], p_limb(K - LO - 3));
        mac_pairs<K, LO + 4, HI>(acc, top, a, b, m);
    } else if constexpr (HI - LO >= 2) {
        if constexpr (opens)
            mac4_new(acc, top, a.l[LO], b.l[K - LO], m[LO], p_limb(K - LO), a.l[LO + 1], b.l[K - LO - 1],
                     m[LO + 1], p_limb(K - LO - 1));
        else
            mac4(acc, top, a.l[LO], b.l[K - LO], m[LO], p_limb(K - LO), a.l[LO + 1], b.l[K - LO - 1],
                 m[LO + 1], p_limb(K - LO - 1));
        mac_pairs<K, LO + 2, HI>(acc, top, a, b, m);
    } else if constexpr (HI - LO == 1) {
        if constexpr (opens)
            mac2_new(acc, top, a.l[LO], b.l[K - LO], m[LO], p_limb(K - LO));
        else
            mac2(acc, top, a.l[LO], b.l[K - LO], m[LO], p_limb(K - LO));
    }
}

// next column: acc = (top : acc) >> 32 (top is restarted by the column's
// first group, or explicitly when a column opens with a single product)
__device__ __forceinline__ void next_column(uint64_t& acc, uint32_t top) {
    acc = (acc >> 32) | ((uint64_t)top << 32);
}

template <int K>
__device__ __forceinline__ void mul_low_column(uint64_t& acc, uint32_t& top, const fp& a, const fp& b,
                                               uint32_t m[12]) {
    if constexpr (K == 0) top = 0;
    mac_pairs<K, 0, K>(acc, top, a, b, m);
    mac1(acc, top, a.l[K], b.l[0]);
    m[K] = (uint32_t)acc * BLS_N0;
    mac1_ps(acc, top, m[K], p_limb(0));
    next_column(acc, top);
    if constexpr (K + 1 < 12) mul_low_column<K + 1>(acc, top, a, b, m);
}

template <int K>
__device__ __forceinline__ void mul_high_column(uint64_t& acc, uint32_t& top, const fp& a, const fp& b,
                                                const uint32_t m[12], uint32_t t[12]) {
    mac_pairs<K, K - 11, 12>(acc, top, a, b, m);
    t[K - 12] = (uint32_t)acc;
    next_column(acc, top);
    if constexpr (K + 1 < 23) mul_high_column<K + 1>(acc, top, a, b, m, t);
}

// Product-scanning (column-wise) Montgomery multiplication with the
// reduction interleaved per column (the "FIPS" ordering of Koc et al.):
// column k gathers a_i*b_(k-i) and m_i*p_(k-i) into a 96-bit accumulator,
// the low word of column k < 12 fixes m_k = lo * (-1/p) mod 2^32, and the
// columns 12..23 are the result words.  288 multiply-adds, each with one
// carry fold, and no per-product register shuffling.  Inputs below 2p give
// a result below 2p (no final subtraction).
__device__ __forceinline__ void fp_mul(fp& r, const fp& a, const fp& b) {
    uint32_t m[12];
    uint32_t t[12];
    uint64_t acc = 0;
    uint32_t top = 0;
    mul_low_column<0>(acc, top, a, b, m);
    mul_high_column<12>(acc, top, a, b, m, t);
    t[11] = (uint32_t)acc;   // result < 2p < 2^382: nothing above word 11
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = t[i];
}

__device__ __forceinline__ void fp_sqr(fp& r, const fp& a) { fp_mul(r, a, a); }

// Sum of two products with one reduction: three products per column index
// (a_i*b_j, c_i*d_j, m_i*p_j), up to 36 per column, so the column sum stays
// below 2^70 and fits the 96-bit accumulator.
#define BLS_MAC3(FIRST)                                         \
    "v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t" FIRST               \
    "v_mad_u64_u32 %0, vcc, %4, %5, %0\n\t"                     \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %6, %7, %0\n\t"                     \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
#define BLS_MAC6(FIRST)                                         \
    BLS_MAC3(FIRST) "\n\t"                                      \
    "v_mad_u64_u32 %0, vcc, %8, %9, %0\n\t"                     \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %10, %11, %0\n\t"                   \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"                  \
    "v_mad_u64_u32 %0, vcc, %12, %13, %0\n\t"                   \
    "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
#define BLS_ADDC_TOP "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
#define BLS_NEW_TOP "v_cndmask_b32_e64 %1, 0, 1, vcc\n\t"

// acc += a0*b0 + c0*d0 + m0*p0
template <bool OPENS>
__device__ __forceinline__ void mac3(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t c0,
                                     uint32_t d0, uint32_t m0, uint32_t p0) {
    if constexpr (OPENS)
        asm(BLS_MAC3(BLS_NEW_TOP)
            : "+v"(acc), "=&v"(top)
            : "v"(a0), "v"(b0), "v"(c0), "v"(d0), "v"(m0), "s"(p0)
            : "vcc");
    else
        asm(BLS_MAC3(BLS_ADDC_TOP)
            : "+v"(acc), "+v"(top)
            : "v"(a0), "v"(b0), "v"(c0), "v"(d0), "v"(m0), "s"(p0)
            : "vcc");
}

// two column indices of the above
template <bool OPENS>
__device__ __forceinline__ void mac6(uint64_t& acc, uint32_t& top, uint32_t a0, uint32_t b0, uint32_t c0,
                                     uint32_t d0, uint32_t m0, uint32_t p0, uint32_t a1, uint32_t b1,
                                     uint32_t c1, uint32_t d1, uint32_t m1, uint32_t p1) {
    if constexpr (OPENS)
        asm(BLS_MAC6(BLS_NEW_TOP)
            : "+v"(acc), "=&v"(top)
            : "v"(a0), "v"(b0), "v"(c0), "v"(d0), "v"(m0), "s"(p0), "v"(a1), "v"(b1), "v"(c1), "v"(d1),
              "v"(m1), "s"(p1)
            : "vcc");
    else
        asm(BLS_MAC6(BLS_ADDC_TOP)
            : "+v"(acc), "+v"(top)
            : "v"(a0), "v"(b0), "v"(c0), "v"(d0), "v"(m0), "s"(p0), "v"(a1), "v"(b1), "v"(c1), "v"(d1),
              "v"(m1), "s"(p1)
            : "vcc");
}

#undef BLS_MAC6
#undef BLS_MAC3
#undef BLS_ADDC_TOP
#undef BLS_NEW_TOP

// the column's triples a_i*b_(k-i) + c_i*d_(k-i) + m_i*p_(k-i), i in [lo, hi)
template <int K, int LO, int HI>
__device__ __forceinline__ void mac_triples(uint64_t& acc, uint32_t& top, const fp& a, const fp& b,
                                            const fp& c, const fp& d, const uint32_t m[12]) {
    constexpr bool opens = LO == 0 || (K >= 12 && LO == K - 11);
    if constexpr (HI - LO >= 2) {
        mac6<opens>(acc, top, a.l[LO], b.l[K - LO], c.l[LO], d.l[K - LO], m[LO], p_limb(K - LO),
                    a.l[LO + 1], b.l[K - LO - 1], c.l[LO + 1], d.l[K - LO - 1], m[LO + 1], p_limb(K - LO - 1));
        mac_triples<K, LO + 2, HI>(acc, top, a, b, c, d, m);
    } else if constexpr (HI - LO == 1) {
        mac3<opens>(acc, top, a.l[LO], b.l[K - LO], c.l[LO], d.l[K - LO], m[LO], p_limb(K - LO));
    }
}

template <int K>
__device__ __forceinline__ void mul_sum_low_column(uint64_t& acc, uint32_t& top, const fp& a, const fp& b,
                                                   const fp& c, const fp& d, uint32_t m[12]) {
    if constexpr (K == 0) top = 0;
    mac_triples<K, 0, K>(acc, top, a, b, c, d, m);
    mac1(acc, top, a.l[K], b.l[0]);
    mac1(acc, top, c.l[K], d.l[0]);
    m[K] = (uint32_t)acc * BLS_N0;
    mac1_ps(acc, top, m[K], p_limb(0));
    next_column(acc, top);
    if constexpr (K + 1 < 12) mul_sum_low_column<K + 1>(acc, top, a, b, c, d, m);
}

template <int K>
__device__ __forceinline__ void mul_sum_high_column(uint64_t& acc, uint32_t& top, const fp& a, const fp& b,
                                                    const fp& c, const fp& d, const uint32_t m[12],
                                                    uint32_t t[12]) {
    mac_triples<K, K - 11, 12>(acc, top, a, b, c, d, m);
    t[K - 12] = (uint32_t)acc;
    next_column(acc, top);
    if constexpr (K + 1 < 23) mul_sum_high_column<K + 1>(acc, top, a, b, c, d, m, t);
}

// r = (a*b + c*d) / R mod p: the columns of fp_mul with both products in
// them and one Montgomery reduction.  Inputs below 2p give a*b + c*d < 8p^2,
// so the result is below 8p^2/R + p < 1.9p: no final subtraction, as in
// fp_mul.
__device__ __forceinline__ void fp_mul_sum(fp& r, const fp& a, const fp& b, const fp& c, const fp& d) {
    uint32_t m[12];
    uint32_t t[12];
    uint64_t acc = 0;
    uint32_t top = 0;
    mul_sum_low_column<0>(acc, top, a, b, c, d, m);
    mul_sum_high_column<12>(acc, top, a, b, c, d, m, t);
    t[11] = (uint32_t)acc;
#pragma unroll
    for (int i = 0; i < 12; i++) r.l[i] = t[i];
}

// a^-1 in Montgomery form, 0 for 0: the divstep inverse of the canonical
// word image c = aR gives c^-1 = a^-1 R^-1, and one product with R^3 mod p
// (c^-1 R^3 / R) brings it to a^-1 R.
__device__ __forceinline__ void fp_inv(fp& r, const fp& a) {
    constexpr uint32_t R3[12] = {0xd94ca1e0, 0xed48ac6b, 0x03a7adf8, 0x315f831e, 0x615e29dd, 0x9a53352a,
                                 0x921e1761, 0x34c04e5e, 0x65724728, 0x2512d435, 0x91755d4d, 0x0aa63460};
    fp c, t, r3;
    fp_canon(c, a);
    divstep::inverse(t.l, c.l);   // below 2p
#pragma unroll
    for (int i = 0; i < 12; i++) r3.l[i] = R3[i];
    fp_mul(r, t, r3);
}

}  // namespace bls
