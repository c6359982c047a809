// Seeded synthetic MSM workloads for benchmarks and the serving loop's load
// generator: uniform scalars of Fr, KZG setup exponents tau^i, the edge-case
// mix of the plain-MSM workload, Montgomery-form blst_fr scalars and Jacobian
// representatives of affine points.  Random source: xorshift64*
// (Vigna), the generator of the reference's tests (test/InsecureRng.h).
#include <stdint.h>
#include <string.h>

#include <vector>

#include "host_ec.h"
#include "host_ec2.h"

namespace {

typedef unsigned __int128 u128;

// group order r of G1 (blst/scalar/fields/Fr.h:127-129)
const uint64_t kR[4] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL,
                        0x73eda753299d7d48ULL};

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed ? seed : 1) {}
    uint64_t next() {
        uint64_t x = s;
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        s = x;
        return x * 2685821657736338717ULL;
    }
};

bool below_r(const uint64_t v[4]) {
    for (int i = 3; i >= 0; i--)
        if (v[i] != kR[i]) return v[i] < kR[i];
    return false;
}

// 4 draws, bit 255 cleared, rejected unless below r
void uniform_fr(Rng& g, uint64_t v[4]) {
    for (;;) {
        for (int i = 0; i < 4; i++) v[i] = g.next();
        v[3] &= 0x7fffffffffffffffULL;
        if (below_r(v)) return;
    }
}

void store32(uint8_t* out, const uint64_t v[4]) {
    for (int i = 0; i < 32; i++) out[i] = (uint8_t)(v[i / 8] >> (8 * (i % 8)));
}

// a*b mod r by a 512-bit product and bitwise reduction (setup-time only)
void mul_mod_r(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
    uint64_t t[8] = {0};
    for (int i = 0; i < 4; i++) {
        uint64_t c = 0;
        for (int j = 0; j < 4; j++) {
            u128 s = (u128)a[i] * b[j] + t[i + j] + c;
            t[i + j] = (uint64_t)s;
            c = (uint64_t)(s >> 64);
        }
        t[i + 4] = c;
    }
    uint64_t rem[5] = {0, 0, 0, 0, 0};
    for (int bit = 511; bit >= 0; bit--) {
        for (int k = 4; k > 0; k--) rem[k] = (rem[k] << 1) | (rem[k - 1] >> 63);
        rem[0] = (rem[0] << 1) | ((t[bit / 64] >> (bit % 64)) & 1);
        bool ge = rem[4] != 0;
        if (!ge) {
            ge = true;
            for (int k = 3; k >= 0; k--)
                if (rem[k] != kR[k]) { ge = rem[k] > kR[k]; break; }
        }
        if (ge) {
            uint64_t br = 0;
            for (int k = 0; k < 4; k++) {
                u128 s = (u128)rem[k] - kR[k] - br;
                rem[k] = (uint64_t)s;
                br = (uint64_t)(s >> 64) & 1;
            }
            rem[4] -= br;
        }
    }
    memcpy(r, rem, 32);
}

// Montgomery multiplication mod r on four 64-bit limbs (CIOS), a * b / 2^256
void mont_mul_r(uint64_t out[4], const uint64_t a[4], const uint64_t b[4]) {
    const uint64_t n0 = 0xfffffffeffffffffULL;   // -r^-1 mod 2^64
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (u128)a[j] * b[i] + t[j];
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        const uint64_t m = t[0] * n0;
        c = ((u128)m * kR[0] + t[0]) >> 64;
        for (int j = 1; j < 4; j++) {
            c += (u128)m * kR[j] + t[j];
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    uint64_t d[4];
    uint64_t borrow = 0;
    for (int j = 0; j < 4; j++) {
        const u128 v = (u128)t[j] - kR[j] - borrow;
        d[j] = (uint64_t)v;
        borrow = (uint64_t)(v >> 64) & 1;
    }
    const bool ge = t[4] != 0 || borrow == 0;
    for (int j = 0; j < 4; j++) out[j] = ge ? d[j] : t[j];
}

}  // namespace

extern "C" {

void blst_hip_synth_scalars(uint8_t* out32, size_t n, uint64_t seed, int low64_pct) {
    Rng g(seed);
    for (size_t i = 0; i < n; i++) {
        uint64_t v[4];
        uniform_fr(g, v);
        if (low64_pct > 0 && (int)(g.next() % 100) < low64_pct) v[1] = v[2] = v[3] = 0;
        store32(out32 + 32 * i, v);
    }
}

void blst_hip_synth_srs_scalars(uint8_t* out32, size_t n, uint64_t seed) {
    Rng g(seed);
    uint64_t tau[4], pw[4] = {1, 0, 0, 0};
    uniform_fr(g, tau);
    for (size_t i = 0; i < n; i++) {
        store32(out32 + 32 * i, pw);
        mul_mod_r(pw, pw, tau);
    }
}

// 1% scalar 0, 1% scalar 1, 1% scalar r-1, 0.5% points replaced by a copy of
// another point; positions are distinct, drawn by a partial Fisher-Yates.
void blst_hip_synth_edge_cases(void* points96, uint8_t* scalars32, size_t n, uint64_t seed) {
    if (n < 2) return;
    uint8_t* pts = static_cast<uint8_t*>(points96);
    Rng g(seed);
    std::vector<size_t> perm(n);
    for (size_t i = 0; i < n; i++) perm[i] = i;
    const size_t n1 = n / 100, n05 = n / 200;
    size_t total = 3 * n1 + n05;
    if (total > n) total = n;
    for (size_t i = 0; i < total; i++) {
        const size_t j = i + (size_t)(g.next() % (uint64_t)(n - i));
        const size_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
    uint64_t rm1[4];
    memcpy(rm1, kR, 32);
    rm1[0] -= 1;
    const uint64_t zero[4] = {0, 0, 0, 0}, one[4] = {1, 0, 0, 0};
    size_t k = 0;
    for (size_t i = 0; i < n1; i++) store32(scalars32 + 32 * perm[k++], zero);
    for (size_t i = 0; i < n1; i++) store32(scalars32 + 32 * perm[k++], one);
    for (size_t i = 0; i < n1; i++) store32(scalars32 + 32 * perm[k++], rm1);
    for (size_t i = 0; i < n05 && k < n; i++, k++) {
        const size_t dst = perm[k];
        size_t src = (size_t)(g.next() % (uint64_t)n);
        if (src == dst) src = (src + 1) % n;
        memcpy(pts + 96 * dst, pts + 96 * src, 96);
    }
}

// blst_fr (Montgomery form, k * 2^256 mod r as four little-endian 64-bit
// limbs) of n canonical 32-byte scalars below r
void blst_hip_synth_fr(uint64_t* out4, const uint8_t* scalars32, size_t n) {
    static const uint64_t kRR[4] = {0xc999e990f3f29c6dULL, 0x2b6cedcb87925c23ULL, 0x05d314967254398fULL,
                                    0x0748d9d99f59ff11ULL};   // 2^512 mod r
    for (size_t i = 0; i < n; i++) {
        uint64_t k[4];
        memcpy(k, scalars32 + 32 * i, 32);
        mont_mul_r(out4 + 4 * i, k, kRR);
    }
}

// Jacobian representatives (x l^2, y l^3, l) of n affine points, l drawn per
// point from the seeded stream (five random 64-bit words: nonzero and below
// p); all-zero affine points give the all-zero Jacobian point
void blst_hip_synth_jacobian(void* points144, const void* affine96, size_t n, uint64_t seed) {
    Rng g(seed);
    const hostec::fp6* a = static_cast<const hostec::fp6*>(affine96);
    hostec::jac6* out = static_cast<hostec::jac6*>(points144);
    for (size_t i = 0; i < n; i++) {
        const hostec::fp6& x = a[2 * i];
        const hostec::fp6& y = a[2 * i + 1];
        hostec::fp6 l;
        for (int w = 0; w < 5; w++) l.l[w] = g.next() | 1;
        l.l[5] = 0;
        if (hostec::is_zero(x) && hostec::is_zero(y)) {
            memset(&out[i], 0, sizeof(out[i]));
            continue;
        }
        hostec::fp6 l2, l3;
        hostec::mul(l2, l, l);
        hostec::mul(l3, l2, l);
        hostec::mul(out[i].x, x, l2);
        hostec::mul(out[i].y, y, l3);
        out[i].z = l;
    }
}

// G2 bases P_i = A + i B, A = [a] G2 and B = [b] G2 for a, b uniform in Fr
// from the seeded stream: one addition per point over Fp2 and one batched
// normalisation (Montgomery's trick over ZZ * ZZZ); affine out, 24 limbs each
void blst_hip_synth_p2_points(void* points192, size_t n, uint64_t seed) {
    using hostec2::f2;
    using hostec2::xyzz12;
    if (n == 0) return;
    // G2 generator, Montgomery form (blst/scalar/curve/point.cpp:114-138)
    static const uint64_t kG2[24] = {
        0xf5f28fa202940a10ULL, 0xb3f5fb2687b4961aULL, 0xa1a893b53e2ae580ULL, 0x9894999d1a3caee9ULL,
        0x6f67b7631863366bULL, 0x058191924350bcd7ULL, 0xa5a9c0759e23f606ULL, 0xaaa0c59dbccd60c3ULL,
        0x3bb17e18e2867806ULL, 0x1b1ab6cc8541b367ULL, 0xc2b6ed0ef2158547ULL, 0x11922a097360edf3ULL,
        0x4c730af860494c4aULL, 0x597cfa1f5e369c5aULL, 0xe7e6856caa0a635aULL, 0xbbefb5e96e0d495fULL,
        0x07d3a975f0ef25a2ULL, 0x0083fd8e7e80dae5ULL, 0xadc0fc92df64b05dULL, 0x18aa270a2b1461dcULL,
        0x86adac6a3be4eba0ULL, 0x79495c4ec93da33aULL, 0xe7175850a43ccaedULL, 0x0b2bc2a163de1bf2ULL};
    xyzz12 g;
    memcpy(&g.x, kG2, 96);
    memcpy(&g.y, kG2 + 12, 96);
    memset(&g.zz, 0, sizeof(g.zz));
    g.zz.c0 = hostec::one();
    g.zzz = g.zz;
    Rng rng(seed);
    auto mult = [&](xyzz12& out) {
        uint64_t k[4];
        uniform_fr(rng, k);
        xyzz12 acc;
        memset(&acc, 0, sizeof(acc));
        for (int b = 255; b >= 0; b--) {
            hostec2::dbl(acc, acc);
            if ((k[b / 64] >> (b % 64)) & 1) hostec2::addp(acc, acc, g);
        }
        out = acc;
    };
    xyzz12 A, B;
    mult(A);
    mult(B);
    std::vector<xyzz12> pts(n);
    pts[0] = A;
    for (size_t i = 1; i < n; i++) hostec2::addp(pts[i], pts[i - 1], B);
    // t_i = ZZ_i ZZZ_i; prefix products, one inversion, walk back
    std::vector<f2> pre(n);
    f2 acc;
    memset(&acc, 0, sizeof(acc));
    acc.c0 = hostec::one();
    for (size_t i = 0; i < n; i++) {
        pre[i] = acc;
        if (!hostec2::is_inf(pts[i])) {
            f2 t;
            hostec2::mul(t, pts[i].zz, pts[i].zzz);
            hostec2::mul(acc, acc, t);
        }
    }
    // 1 / (c0 + c1 i) = (c0 - c1 i) / (c0^2 + c1^2)
    hostec::fp6 nrm, t0, ninv;
    hostec::mul(nrm, acc.c0, acc.c0);
    hostec::mul(t0, acc.c1, acc.c1);
    hostec::add(nrm, nrm, t0);
    hostec::inv(ninv, nrm);
    f2 inv;
    hostec::mul(inv.c0, acc.c0, ninv);
    hostec::mul(t0, acc.c1, ninv);
    hostec::fp6 zero{};
    hostec::sub(inv.c1, zero, t0);
    uint64_t* out = static_cast<uint64_t*>(points192);
    for (size_t i = n; i-- > 0;) {
        uint64_t* o = out + 24 * i;
        if (hostec2::is_inf(pts[i])) {
            memset(o, 0, 192);
            continue;
        }
        f2 it, t, izz, izzz, x, y;
        hostec2::mul(it, inv, pre[i]);                 // 1 / (ZZ ZZZ)
        hostec2::mul(t, pts[i].zz, pts[i].zzz);
        hostec2::mul(inv, inv, t);
        hostec2::mul(izz, it, pts[i].zzz);             // 1 / ZZ
        hostec2::mul(izzz, it, pts[i].zz);             // 1 / ZZZ
        hostec2::mul(x, pts[i].x, izz);
        hostec2::mul(y, pts[i].y, izzz);
        memcpy(o, &x, 96);
        memcpy(o + 12, &y, 96);
    }
}

}  // extern "C"
