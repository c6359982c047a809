// Seeded synthetic MSM workloads for benchmarks and the serving loop's load
// generator: uniform scalars of Fr, KZG setup exponents tau^i, the edge-case
// mix of the plain-MSM workload, Montgomery-form blst_fr scalars and Jacobian
// representatives of affine points.  Random source: xorshift64*
// (Vigna), the generator of the reference's tests (test/InsecureRng.h).
#include <stdint.h>
#include <string.h>

#include <vector>

#include "host_ec.h"

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

}  // extern "C"
