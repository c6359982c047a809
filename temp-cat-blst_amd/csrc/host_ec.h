// Host-side BLS12-381 G1 arithmetic for the latency-bound tail of an MSM:
// folding the per-window sums W_j into sum_j 2^(c*j) W_j (c*D doublings in a
// row, which a single CPU core finishes in tens of microseconds while one GPU
// lane would need milliseconds), and the 144-byte blst_p1 output.
//
// Montgomery arithmetic on six 64-bit limbs, R = 2^384, the layout of
// blst_fp (bindings/blst_types.h:95-152).  Formulas: add-2008-s and
// dbl-2008-s-1 in XYZZ coordinates, Pointxyzz::to_Jacobian
// (blst/scalar/curve/point.cpp:1939-2001, 2109-2114).
#pragma once
#include <stdint.h>
#include <string.h>
#if defined(__x86_64__) && !defined(__HIP_DEVICE_COMPILE__)
#include <cpuid.h>
#define HOSTEC_X86_ADX 1
#endif

namespace hostec {

typedef unsigned __int128 u128;

struct fp6 { uint64_t l[6]; };
struct xyzz6 { fp6 x, y, zz, zzz; };
struct jac6 { fp6 x, y, z; };

static const uint64_t kP[6] = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
                               0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
static const uint64_t kN0 = 0x89f3fffcfffcfffdULL;

inline bool is_zero(const fp6& a) {
    return (a.l[0] | a.l[1] | a.l[2] | a.l[3] | a.l[4] | a.l[5]) == 0;
}

inline void sub_p_if_ge(fp6& r, const uint64_t t[7]) {
    uint64_t d[6], br = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)t[j] - kP[j] - br;
        d[j] = (uint64_t)s;
        br = (uint64_t)(s >> 64) & 1;
    }
    const bool keep = t[6] < br;
    for (int j = 0; j < 6; j++) r.l[j] = keep ? t[j] : d[j];
}

// Column accumulator of the product below: three words, bits 0..191.
struct acc192 { uint64_t lo, hi, ov; };

// s += x y.  The high word of a 64 x 64-bit product is at most 2^64 - 2, so
// the carry of the low word added to it cannot wrap.
inline void mac(acc192& s, uint64_t x, uint64_t y) {
    const u128 p = (u128)x * y;
    unsigned long long t;
    const uint64_t hi = (uint64_t)(p >> 64) + __builtin_uaddll_overflow(s.lo, (uint64_t)p, &t);
    s.lo = t;
    s.ov += __builtin_uaddll_overflow(s.hi, hi, &t);
    s.hi = t;
}

// Montgomery product by product scanning with the reduction interleaved per
// column (the order of the device product in fp.hip.h, on 64-bit words):
// column k gathers a_i b_(k-i) and m_i p_(k-i) in the 192-bit accumulator,
// the low word of column k < 6 fixes m_k = lo (-1/p) mod 2^64 (which clears
// it), and columns 6..11 are the result words, less p if needed.  Inputs
// below 2p give a canonical result (a b + m p < 4p^2 + 2^384 p < 2^384 2p).
inline void mul_portable(fp6& r, const fp6& a, const fp6& b) {
    uint64_t m[6], t[7];
    acc192 s = {0, 0, 0};
    for (int k = 0; k < 6; k++) {
        for (int i = 0; i < k; i++) {
            mac(s, a.l[i], b.l[k - i]);
            mac(s, m[i], kP[k - i]);
        }
        mac(s, a.l[k], b.l[0]);
        m[k] = s.lo * kN0;
        mac(s, m[k], kP[0]);
        s = {s.hi, s.ov, 0};
    }
    for (int k = 6; k < 11; k++) {
        for (int i = k - 5; i < 6; i++) {
            mac(s, a.l[i], b.l[k - i]);
            mac(s, m[i], kP[k - i]);
        }
        t[k - 6] = s.lo;
        s = {s.hi, s.ov, 0};
    }
    t[5] = s.lo;
    t[6] = s.hi;
    sub_p_if_ge(r, t);
}

#ifdef HOSTEC_X86_ADX
// The same product on x86-64 with BMI2 and ADX, operand scanning (CIOS) in
// six rounds over a window of seven words L0..L6.  Round i adds a b_i with
// mulx, the low words along the CF chain (adcx) and the high words along the
// OF chain (adox), so the two carry chains run side by side; then it adds
// m p with m = L0 (-1/p) mod 2^64 the same way, which clears L0.  The next
// round's window is L1..L6 and the cleared word, so the rounds name the seven
// registers (%0..%6) rotated by one.  p's top word is below 2^62, so with
// inputs below 2p the running value stays below 2^383 and the window never
// needs an eighth word; the result is below a b / 2^384 + p < 1.4p and one
// subtraction of p makes it canonical, as above.
// Operands: %7, %8 the product halves, %9 a, %10 b, %11 p, %12 -1/p mod 2^64.
#define HOSTEC_MAC(OFF, SRC, LO, HI)       \
    "mulxq " #OFF "(%" #SRC "), %7, %8\n\t" \
    "adcxq %7, %" #LO "\n\t"                \
    "adoxq %8, %" #HI "\n\t"
#define HOSTEC_ROUND(I, L0, L1, L2, L3, L4, L5, L6)                                            \
    "movq " #I "*8(%10), %%rdx\n\t"                                                            \
    "xorq %" #L6 ", %" #L6 "\n\t"                                                              \
    HOSTEC_MAC(0, 9, L0, L1) HOSTEC_MAC(8, 9, L1, L2) HOSTEC_MAC(16, 9, L2, L3)                  \
    HOSTEC_MAC(24, 9, L3, L4) HOSTEC_MAC(32, 9, L4, L5) HOSTEC_MAC(40, 9, L5, L6)                \
    "adcq $0, %" #L6 "\n\t"                                                                    \
    "movq %" #L0 ", %%rdx\n\t"                                                                 \
    "imulq %12, %%rdx\n\t"                                                                     \
    "xorq %7, %7\n\t"                                                                          \
    HOSTEC_MAC(0, 11, L0, L1) HOSTEC_MAC(8, 11, L1, L2) HOSTEC_MAC(16, 11, L2, L3)               \
    HOSTEC_MAC(24, 11, L3, L4) HOSTEC_MAC(32, 11, L4, L5) HOSTEC_MAC(40, 11, L5, L6)             \
    "adcq $0, %" #L6 "\n\t"

inline void mul_adx(fp6& r, const fp6& a, const fp6& b) {
    uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0, w6, lo, hi;
    asm(HOSTEC_ROUND(0, 0, 1, 2, 3, 4, 5, 6) HOSTEC_ROUND(1, 1, 2, 3, 4, 5, 6, 0)
        HOSTEC_ROUND(2, 2, 3, 4, 5, 6, 0, 1) HOSTEC_ROUND(3, 3, 4, 5, 6, 0, 1, 2)
        HOSTEC_ROUND(4, 4, 5, 6, 0, 1, 2, 3) HOSTEC_ROUND(5, 5, 6, 0, 1, 2, 3, 4)
        : "+&r"(w0), "+&r"(w1), "+&r"(w2), "+&r"(w3), "+&r"(w4), "+&r"(w5), "=&r"(w6), "=&r"(lo), "=&r"(hi)
        : "r"(a.l), "r"(b.l), "r"(kP), "m"(kN0)
        : "rdx", "cc", "memory");
    const uint64_t t[7] = {w6, w0, w1, w2, w3, w4, 0};
    sub_p_if_ge(r, t);
}
#undef HOSTEC_ROUND
#undef HOSTEC_MAC

// whether this CPU has BMI2 (mulx) and ADX (adcx, adox): CPUID leaf 7,
// EBX bits 8 and 19
inline bool have_adx() {
    static const bool v = [] {
        unsigned a, b, c, d;
        return __get_cpuid_count(7, 0, &a, &b, &c, &d) && ((b >> 8) & 1) && ((b >> 19) & 1);
    }();
    return v;
}
#endif

// r = a b 2^-384 mod p, canonical for inputs below 2p
inline void mul(fp6& r, const fp6& a, const fp6& b) {
#ifdef HOSTEC_X86_ADX
    if (have_adx()) {
        mul_adx(r, a, b);
        return;
    }
#endif
    mul_portable(r, a, b);
}

inline void add(fp6& r, const fp6& a, const fp6& b) {
    uint64_t t[7], c = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)a.l[j] + b.l[j] + c;
        t[j] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    t[6] = c;
    sub_p_if_ge(r, t);
}

inline void sub(fp6& r, const fp6& a, const fp6& b) {
    uint64_t d[6], br = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)a.l[j] - b.l[j] - br;
        d[j] = (uint64_t)s;
        br = (uint64_t)(s >> 64) & 1;
    }
    const uint64_t mask = 0 - br;
    uint64_t c = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)d[j] + (kP[j] & mask) + c;
        r.l[j] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
}

// Montgomery one, R mod p
static const uint64_t kOne[6] = {0x760900000002fffdULL, 0xebf4000bc40c0002ULL, 0x5f48985753c758baULL,
                                 0x77ce585370525745ULL, 0x5c071a97a256ec6dULL, 0x15f65ec3fa80e493ULL};

inline fp6 one() { fp6 r; memcpy(r.l, kOne, 48); return r; }

// a^(p-2) by left-to-right square-and-multiply; 0 maps to 0
inline void inv(fp6& r, const fp6& a) {
    uint64_t e[6];
    memcpy(e, kP, 48);
    e[0] -= 2;
    fp6 acc = one();
    for (int bit = 380; bit >= 0; bit--) {
        mul(acc, acc, acc);
        if ((e[bit / 64] >> (bit % 64)) & 1) mul(acc, acc, a);
    }
    r = acc;
}

inline void set_inf(xyzz6& p) { memset(&p, 0, sizeof(p)); }
inline bool is_inf(const xyzz6& p) { return is_zero(p.zz); }

inline void dbl(xyzz6& r, const xyzz6& p) {
    if (is_inf(p)) { r = p; return; }
    fp6 U, V, W, S, M, t;
    add(U, p.y, p.y);
    mul(V, U, U);
    mul(W, U, V);
    mul(S, p.x, V);
    mul(M, p.x, p.x);
    add(t, M, M);
    add(M, M, t);
    xyzz6 o;
    mul(o.x, M, M);
    add(t, S, S);
    sub(o.x, o.x, t);
    sub(t, S, o.x);
    mul(t, t, M);
    mul(o.y, W, p.y);
    sub(o.y, t, o.y);
    mul(o.zz, V, p.zz);
    mul(o.zzz, W, p.zzz);
    r = o;
}

inline void addp(xyzz6& r, const xyzz6& p, const xyzz6& q) {
    if (is_inf(q)) { r = p; return; }
    if (is_inf(p)) { r = q; return; }
    fp6 U1, S1, P, R;
    mul(U1, p.x, q.zz);
    mul(S1, p.y, q.zzz);
    mul(P, q.x, p.zz);
    mul(R, q.y, p.zzz);
    sub(P, P, U1);
    sub(R, R, S1);
    if (is_zero(P)) {
        if (is_zero(R)) dbl(r, p);
        else set_inf(r);
        return;
    }
    fp6 PP, PPP, Q, t;
    xyzz6 o;
    mul(PP, P, P);
    mul(PPP, PP, P);
    mul(Q, U1, PP);
    mul(t, R, R);
    sub(t, t, PPP);
    sub(t, t, Q);
    sub(o.x, t, Q);
    sub(Q, Q, o.x);
    mul(Q, Q, R);
    mul(t, S1, PPP);
    sub(o.y, Q, t);
    mul(t, p.zz, q.zz);
    mul(o.zz, t, PP);
    mul(t, p.zzz, q.zzz);
    mul(o.zzz, t, PPP);
    r = o;
}

// Pointxyzz::to_Jacobian; infinity maps to the all-zero blst_p1 the
// reference's early returns produce (c_api.cpp:1390-1408).
inline void to_jac(jac6& r, const xyzz6& p) {
    if (is_inf(p)) { memset(&r, 0, sizeof(r)); return; }
    mul(r.x, p.x, p.zz);
    mul(r.y, p.y, p.zzz);
    r.z = p.zz;
}

// XYZZ -> affine for n points with one inversion (Montgomery's trick over
// ZZ*ZZZ = Z^5); infinity becomes the all-zero affine pair.
struct aff6 { fp6 x, y; };
inline void batch_to_affine(aff6* out, const xyzz6* in, size_t n) {
    fp6* pre = new fp6[n + 1];
    fp6 acc = one();
    for (size_t i = 0; i < n; i++) {
        pre[i] = acc;
        if (!is_inf(in[i])) {
            fp6 t;
            mul(t, in[i].zz, in[i].zzz);
            mul(acc, acc, t);
        }
    }
    fp6 ia;
    inv(ia, acc);
    for (size_t i = n; i-- > 0;) {
        if (is_inf(in[i])) { memset(&out[i], 0, sizeof(aff6)); continue; }
        fp6 t, it, izz, izzz;
        mul(t, in[i].zz, in[i].zzz);
        mul(it, ia, pre[i]);          // 1 / (ZZ*ZZZ)
        mul(ia, ia, t);
        mul(izz, it, in[i].zzz);      // 1 / ZZ
        mul(izzz, it, in[i].zz);      // 1 / ZZZ
        mul(out[i].x, in[i].x, izz);
        mul(out[i].y, in[i].y, izzz);
    }
    delete[] pre;
}

// ZCash compressed encoding of a Jacobian point (the blst_p1_compress
// format, blst/scalar/curve/point.cpp:817-902): 48 big-endian bytes of x,
// 0x80 = compressed, 0x40 = infinity, 0x20 = y > (p-1)/2.
inline void compress(uint8_t out[48], const jac6& p) {
    if (is_zero(p.z)) {
        memset(out, 0, 48);
        out[0] = 0xc0;
        return;
    }
    fp6 zi, zi2, zi3, x, y;
    inv(zi, p.z);
    mul(zi2, zi, zi);
    mul(zi3, zi2, zi);
    mul(x, p.x, zi2);
    mul(y, p.y, zi3);
    fp6 plain1 = {{1, 0, 0, 0, 0, 0}};
    mul(x, x, plain1);   // out of Montgomery form
    mul(y, y, plain1);
    for (int i = 0; i < 48; i++) out[i] = (uint8_t)(x.l[(47 - i) / 8] >> (8 * ((47 - i) % 8)));
    // y > (p-1)/2  <=>  2y > p - 1  <=>  2y >= p (p odd)
    uint64_t d[7], c = 0;
    for (int j = 0; j < 6; j++) {
        d[j] = (y.l[j] << 1) | c;
        c = y.l[j] >> 63;
    }
    d[6] = c;
    uint64_t br = 0;
    for (int j = 0; j < 6; j++) {
        u128 s = (u128)d[j] - kP[j] - br;
        br = (uint64_t)(s >> 64) & 1;
    }
    const bool big = d[6] >= br;
    out[0] |= (uint8_t)(0x80 | (big ? 0x20 : 0));
}

// Jacobian doubling for a = 0 (dbl-2009-l, 1M + 5S), for the long doubling
// runs of the window fold; Z = 0 stays infinity
inline void jdbl(jac6& r, const jac6& p) {
    if (is_zero(p.z)) { r = p; return; }
    fp6 A, B, C, D, E, F, t;
    mul(A, p.x, p.x);
    mul(B, p.y, p.y);
    mul(C, B, B);
    add(t, p.x, B);
    mul(t, t, t);
    sub(t, t, A);
    sub(t, t, C);
    add(D, t, t);            // D = 2((X + B)^2 - A - C) = 4XY^2
    add(E, A, A);
    add(E, E, A);            // E = 3X^2
    mul(F, E, E);
    jac6 o;
    sub(o.x, F, D);
    sub(o.x, o.x, D);        // X3 = E^2 - 2D
    sub(t, D, o.x);
    mul(t, E, t);
    add(C, C, C);
    add(C, C, C);
    add(C, C, C);            // 8C
    sub(o.y, t, C);          // Y3 = E(D - X3) - 8C
    mul(o.z, p.y, p.z);
    add(o.z, o.z, o.z);      // Z3 = 2YZ
    r = o;
}

// 2^bits * p: the doublings in Jacobian coordinates, back to XYZZ
// (ZZ = Z^2, ZZZ = Z^3)
inline void scale_pow2(xyzz6& r, const xyzz6& p, int bits) {
    if (bits == 0 || is_inf(p)) { r = p; return; }
    jac6 j;
    to_jac(j, p);
    for (int k = 0; k < bits; k++) jdbl(j, j);
    r.x = j.x;
    r.y = j.y;
    mul(r.zz, j.z, j.z);
    mul(r.zzz, r.zz, j.z);
}

// sum_j 2^(c*j) w[j], Horner from the top window
inline void fold_windows(xyzz6& out, const xyzz6* w, int nwin, int c) {
    xyzz6 acc;
    set_inf(acc);
    for (int j = nwin - 1; j >= 0; j--) {
        if (j != nwin - 1) scale_pow2(acc, acc, c);
        addp(acc, acc, w[j]);
    }
    out = acc;
}

}  // namespace hostec
