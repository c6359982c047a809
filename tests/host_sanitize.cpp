// Host code of the product checked against the oracle, for the sanitizer
// builds of tests/test_host_sanitizers.py:
//   host_ec.h    G1 window fold, XYZZ -> Jacobian, compression, batch
//                normalisation and inversion (the host tail of every G1 MSM),
//                and its Montgomery product in both forms
//   host_ec2.h   the G2 window fold and XYZZ -> Jacobian
//   fq.hip.h     the G1 accumulation's 28-bit limb field and its bucket sum
//   fp_inv.hip.h the divstep inversion
// built together with the oracle's C sources under AddressSanitizer and
// UndefinedBehaviorSanitizer (mode "asan"), and the oracle's threaded MSMs
// and point generation under ThreadSanitizer (mode "tsan").  Every check
// compares two computations and the process exits with status 1 at the first
// mismatch; the sanitizers abort at the first error they report.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../oracle/blst_oracle.h"
#include "../temp-cat-blst_amd/csrc/host_ec.h"
#include "../temp-cat-blst_amd/csrc/host_ec2.h"
#define BLS_HD inline
#include "../temp-cat-blst_amd/csrc/fp_inv.hip.h"
#include "../temp-cat-blst_amd/csrc/fq.hip.h"

namespace {

uint64_t g_rng = 0x243f6a8885a308d3ULL;
uint64_t next64() {
    uint64_t z = (g_rng += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

int g_checks = 0;
void expect(bool ok, const char* what, int a, int b) {
    g_checks++;
    if (!ok) {
        std::printf("MISMATCH %s (%d, %d)\n", what, a, b);
        std::exit(1);
    }
}

// p as six 64-bit limbs, and Montgomery one
const uint64_t kP[6] = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
                        0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};

// a uniform value below p (rejection sampling on 381 bits)
void rand_fp(uint64_t r[6]) {
    for (;;) {
        for (int i = 0; i < 6; i++) r[i] = next64();
        r[5] &= (1ULL << 61) - 1;
        int i = 5;
        while (i >= 0 && r[i] == kP[i]) i--;
        if (i < 0 || r[i] < kP[i]) return;
    }
}

// scalars of nbits bits, ceil(nbits / 8) bytes each, bits above nbits clear
std::vector<uint8_t> scalars(size_t n, size_t nbits, uint64_t seed) {
    const size_t nb = (nbits + 7) / 8;
    std::vector<uint8_t> s32(32 * n + 32), out(nb * n);
    or_gen_scalars(s32.data(), n, seed, 0);
    for (size_t i = 0; i < n; i++) {
        std::memcpy(&out[nb * i], &s32[32 * i], nb);
        if (nbits % 8) out[nb * i + nb - 1] &= (uint8_t)((1u << (nbits % 8)) - 1);
    }
    return out;
}

void jac_to_xyzz(hostec::xyzz6& r, const uint64_t* j) {
    hostec::fp6 z;
    std::memcpy(z.l, j + 12, 48);
    if (hostec::is_zero(z)) {
        hostec::set_inf(r);
        return;
    }
    std::memcpy(r.x.l, j, 48);
    std::memcpy(r.y.l, j + 6, 48);
    hostec::mul(r.zz, z, z);
    hostec::mul(r.zzz, r.zz, z);
}

// Sum_j 2^(c j) tile_j by the product's fold, compressed, against the
// oracle's Pippenger MSM; some scalars zero or one, some points repeated or
// negated
void check_g1_fold() {
    const size_t ns[] = {1, 2, 7, 33};
    const size_t nbitss[] = {1, 8, 64, 255, 256};
    const size_t cs[] = {2, 5, 13};
    for (size_t n : ns) {
        std::vector<uint64_t> pts(12 * n);
        or_gen_points(pts.data(), n, 100 + n, 1);
        if (n > 3) {
            std::memcpy(&pts[12], &pts[0], 96);                          // repeated point
            std::memcpy(&pts[24], &pts[0], 48);                          // -P
            or_fp_neg(&pts[30], &pts[6]);
        }
        for (size_t nbits : nbitss) {
            std::vector<uint8_t> sc = scalars(n, nbits, 7 * n + nbits);
            const size_t nb = (nbits + 7) / 8;
            if (n > 5) {
                std::memset(&sc[nb * 4], 0, nb);                         // zero scalar
                std::memset(&sc[nb * 5], 0, nb);
                sc[nb * 5] = 1;                                          // scalar one
            }
            uint64_t ref[18];
            uint8_t want[48], got[48];
            or_msm_pippenger(ref, pts.data(), sc.data(), n, nbits, 0);
            or_p1_compress(want, ref);
            for (size_t c : cs) {
                const int nwin = (int)(nbits / c) + 1;
                std::vector<hostec::xyzz6> w(nwin);
                for (int j = 0; j < nwin; j++) {
                    uint64_t t[18];
                    or_msm_tile(t, pts.data(), sc.data(), n, nbits, (size_t)j * c, c);
                    jac_to_xyzz(w[j], t);
                }
                hostec::xyzz6 r;
                hostec::fold_windows(r, w.data(), nwin, (int)c);
                hostec::jac6 jr;
                hostec::to_jac(jr, r);
                hostec::compress(got, jr);
                expect(std::memcmp(got, want, 48) == 0, "g1 fold", (int)n, (int)(nbits * 100 + c));
            }
        }
    }
}

// the host Montgomery product: the portable form against the oracle's
// (inputs below p) and, where the CPU has BMI2 and ADX, the mulx/adcx/adox
// form against the portable one on inputs below 2p, edge values included
void check_host_products() {
    const uint64_t twop_m1[6] = {0x73fdffffffff5555ULL, 0x3d57fffd62a7ffffULL, 0xce61a541ed61ec48ULL,
                                 0xc8ee9709e70a257eULL, 0x96374f6c869759aeULL, 0x340223d472ffcd34ULL};
    auto edge = [&](uint64_t r[6], int k) {
        std::memset(r, 0, 48);
        if (k == 1) r[0] = 1;
        if (k == 2) { std::memcpy(r, kP, 48); r[0] -= 1; }
        if (k == 3) std::memcpy(r, kP, 48);
        if (k == 4) std::memcpy(r, twop_m1, 48);
    };
    for (int t = 0; t < 20000; t++) {
        uint64_t a[6], b[6], want[6];
        rand_fp(a);
        rand_fp(b);
        if (t < 25) {
            edge(a, t % 5);
            edge(b, t / 5);
        }
        hostec::fp6 x, y, r1, r2;
        std::memcpy(x.l, a, 48);
        std::memcpy(y.l, b, 48);
        hostec::mul_portable(r1, x, y);
        const bool below_p = t >= 25 || (t % 5 < 3 && t / 5 < 3);
        if (below_p) {
            or_fp_mul(want, a, b);
            expect(std::memcmp(r1.l, want, 48) == 0, "host product", t, 0);
        }
#ifdef HOSTEC_X86_ADX
        if (hostec::have_adx()) {
            if (t >= 25 && (t & 1)) {   // lift into [p, 2p) where that stays below 2p
                uint64_t c = 0;
                for (int i = 0; i < 6; i++) {
                    const unsigned __int128 v = (unsigned __int128)x.l[i] + kP[i] + c;
                    x.l[i] = (uint64_t)v;
                    c = (uint64_t)(v >> 64);
                }
                hostec::mul_portable(r1, x, y);
            }
            hostec::mul_adx(r2, x, y);
            expect(std::memcmp(r1.l, r2.l, 48) == 0, "host product, adx", t, 0);
        }
#endif
    }
}

// batch normalisation and the inversion of host_ec.h against the oracle
void check_g1_affine_and_inverse() {
    const size_t n = 9;
    std::vector<uint64_t> jac(18 * n, 0), want(12 * n), aff(12);
    uint64_t g[12], g18[18];
    or_p1_generator(g);
    or_p1_from_affine(g18, g);
    for (size_t i = 0; i < n; i++) {
        if (i == 3) continue;   // an infinity point
        uint8_t k[32];
        for (int b = 0; b < 32; b++) k[b] = (uint8_t)next64();
        k[31] &= 0x3f;
        or_p1_mult(&jac[18 * i], g18, k, 255);
    }
    or_p1s_to_affine(want.data(), jac.data(), n);
    std::vector<hostec::xyzz6> in(n);
    std::vector<hostec::aff6> out(n);
    for (size_t i = 0; i < n; i++) jac_to_xyzz(in[i], &jac[18 * i]);
    hostec::batch_to_affine(out.data(), in.data(), n);
    for (size_t i = 0; i < n; i++)
        expect(std::memcmp(&out[i], &want[12 * i], 96) == 0, "batch_to_affine", (int)i, 0);
    for (int t = 0; t < 64; t++) {
        uint64_t a[6], r[6];
        rand_fp(a);
        if (t == 0) std::memset(a, 0, 48);
        or_fp_inv(r, a);
        hostec::fp6 x, y;
        std::memcpy(x.l, a, 48);
        hostec::inv(y, x);
        expect(std::memcmp(y.l, r, 48) == 0, "host inverse", t, 0);
    }
}

// the G2 fold against the oracle's G2 Pippenger MSM (compressed)
void check_g2_fold() {
    const size_t ns[] = {1, 5};
    const size_t nbitss[] = {7, 255};
    for (size_t n : ns) {
        std::vector<uint64_t> pts(24 * n);
        or_gen_p2_points(pts.data(), n, 300 + n);
        for (size_t nbits : nbitss) {
            std::vector<uint8_t> sc = scalars(n, nbits, 11 * n + nbits);
            uint64_t ref[36];
            uint8_t want[96], got[96];
            or_msm_p2_pippenger(ref, pts.data(), sc.data(), n, nbits, 0);
            or_p2_compress(want, ref);
            const size_t c = 4;
            const int nwin = (int)(nbits / c) + 1;
            std::vector<hostec2::xyzz12> w(nwin);
            for (int j = 0; j < nwin; j++) {
                uint64_t t[36];
                or_msm_p2_tile(t, pts.data(), sc.data(), n, nbits, (size_t)j * c, c);
                hostec2::f2 z;
                std::memcpy(&z, t + 24, sizeof(z));
                if (hostec2::is_zero(z)) {
                    std::memset(&w[j], 0, sizeof(w[j]));
                    continue;
                }
                std::memcpy(&w[j].x, t, sizeof(hostec2::f2));
                std::memcpy(&w[j].y, t + 12, sizeof(hostec2::f2));
                hostec2::mul(w[j].zz, z, z);
                hostec2::mul(w[j].zzz, w[j].zz, z);
            }
            hostec2::xyzz12 r;
            hostec2::fold_windows(r, w.data(), nwin, (int)c);
            hostec2::jac12 jr;
            hostec2::to_jac(jr, r);
            uint64_t j36[36];
            static_assert(sizeof(jr) == sizeof(j36), "blst_p2 layout");
            std::memcpy(j36, &jr, sizeof(j36));
            or_p2_compress(got, j36);
            expect(std::memcmp(got, want, 96) == 0, "g2 fold", (int)n, (int)nbits);
        }
    }
}

// canonical 32-bit word image (below p) of a value below 2p in the word
// layout, through a Montgomery product by one
void canon(uint64_t r[6], const uint32_t w[12]) {
    hostec::fp6 a, c;
    std::memcpy(a.l, w, 48);
    hostec::mul(c, a, hostec::one());
    std::memcpy(r, c.l, 48);
}

// the divstep inversion (plain integers) against the oracle's inverse
void check_divstep() {
    for (int t = 0; t < 200; t++) {
        uint64_t x[6];
        if (t < 8) {
            std::memset(x, 0, 48);
            if (t == 1) x[0] = 1;
            if (t == 2) { std::memcpy(x, kP, 48); x[0] -= 1; }
            if (t >= 3) x[t - 3] = 1ULL << (t * 7);
        } else {
            rand_fp(x);
        }
        uint32_t in[12], out[12];
        std::memcpy(in, x, 48);
        bls::divstep::inverse(out, in);
        uint64_t xm[6], im[6], want[6], gc[6];
        or_fp_to_mont(xm, x);
        or_fp_inv(im, xm);
        or_fp_from_mont(want, im);
        canon(gc, out);                  // the result is below 2p
        expect(std::memcmp(gc, want, 48) == 0, "divstep inverse", t, 0);
    }
}

// fq products against the oracle's: for canonical words x, y (R = 2^384
// images) the words of the fq product of their R' = 2^392 images, read as
// 32-bit words, are the oracle's x y R^-1 times 2^8 (the crossing rules of
// fq.hip.h); likewise squares and sums of two products
void check_fq_products() {
    for (int t = 0; t < 300; t++) {
        uint64_t a[6], b[6], c[6], d[6];
        rand_fp(a);
        rand_fp(b);
        rand_fp(c);
        rand_fp(d);
        // exact shifted images, then normalized below 1.2p by a product
        // with one (the operand bounds fq_mul states)
        bls::fq A, B, C, D, R, one;
        bls::fq_one(one);
        bls::fq_from_words<8>(A, reinterpret_cast<const uint32_t*>(a));
        bls::fq_from_words<8>(B, reinterpret_cast<const uint32_t*>(b));
        bls::fq_from_words<8>(C, reinterpret_cast<const uint32_t*>(c));
        bls::fq_from_words<8>(D, reinterpret_cast<const uint32_t*>(d));
        bls::fq_mul(A, A, one);
        bls::fq_mul(B, B, one);
        bls::fq_mul(C, C, one);
        bls::fq_mul(D, D, one);
        for (int op = 0; op < 3; op++) {
            uint64_t m[6], e[6];
            if (op == 0) {
                bls::fq_mul(R, A, B);
                or_fp_mul(m, a, b);
            } else if (op == 1) {
                bls::fq_sqr(R, A);
                or_fp_mul(m, a, a);
            } else {
                bls::fq_mul_sum(R, A, B, C, D);
                uint64_t m2[6];
                or_fp_mul(m, a, b);
                or_fp_mul(m2, c, d);
                or_fp_add(m, m, m2);
            }
            for (int k = 0; k < 8; k++) or_fp_add(m, m, m);
            uint32_t w[12];
            bls::fq_to_words(w, R);
            canon(e, w);
            expect(std::memcmp(e, m, 48) == 0, "fq product", t, op);
        }
    }
}

// the accumulation's bucket sum in fq (madd-2008-s with the doubling and
// vanishing cases) against the oracle's Jacobian additions
void check_fq_bucket_sums() {
    const uint32_t P2[12] = {0xffff5556u, 0x73fdffffu, 0x62a7ffffu, 0x3d57fffdu, 0xed61ec48u, 0xce61a541u,
                             0xe70a257eu, 0xc8ee9709u, 0x869759aeu, 0x96374f6cu, 0x72ffcd34u, 0x340223d4u};
    const size_t n = 12;
    std::vector<uint64_t> pts(12 * n);
    or_gen_points(pts.data(), n, 77, 1);
    for (int trial = 0; trial < 40; trial++) {
        const int len = 1 + (int)(next64() % 10);
        bls::fq_xyzz acc;
        bls::fq_xyzz_set_inf(acc);
        uint64_t ref[18] = {0};
        for (int k = 0; k < len; k++) {
            const size_t i = next64() % 4;            // few distinct points: doublings and cancellations
            const bool neg = (next64() & 1) != 0;
            uint32_t x[12], y[12];
            std::memcpy(x, &pts[12 * i], 48);
            std::memcpy(y, &pts[12 * i + 6], 48);
            uint64_t aff[12];
            std::memcpy(aff, &pts[12 * i], 96);
            if (neg) {
                uint64_t borrow = 0;
                for (int j = 0; j < 12; j++) {
                    const uint64_t v = (uint64_t)P2[j] - y[j] - borrow;
                    y[j] = (uint32_t)v;
                    borrow = (v >> 63) & 1;
                }
                or_fp_neg(aff + 6, aff + 6);
            }
            bls::fq fx, fy16;
            bls::fq_from_words<8>(fx, x);
            bls::fq_from_words<4>(fy16, y);
            bls::fq_xyzz_madd(acc, fx, fy16);
            uint64_t q[18];
            or_p1_from_affine(q, aff);
            or_p1_add(ref, ref, q);
        }
        uint32_t w[48];
        bls::fq_xyzz_to_words(w, w + 12, w + 24, w + 36, acc);
        hostec::xyzz6 s;
        uint64_t cx[6];
        canon(cx, w);
        std::memcpy(s.x.l, cx, 48);
        canon(cx, w + 12);
        std::memcpy(s.y.l, cx, 48);
        canon(cx, w + 24);
        std::memcpy(s.zz.l, cx, 48);
        canon(cx, w + 36);
        std::memcpy(s.zzz.l, cx, 48);
        hostec::jac6 j;
        hostec::to_jac(j, s);
        uint8_t got[48], want[48];
        hostec::compress(got, j);
        or_p1_compress(want, ref);
        expect(std::memcmp(got, want, 48) == 0, "fq bucket sum", trial, len);
    }
}

// the oracle's threaded paths against their one-thread forms
void check_threads() {
    const size_t n = 300;
    std::vector<uint64_t> p1(12 * n), p4(12 * n);
    or_gen_points(p1.data(), n, 9, 1);
    or_gen_points(p4.data(), n, 9, 4);
    expect(p1 == p4, "threaded point generation", (int)n, 0);
    std::vector<uint8_t> sc = scalars(n, 255, 5);
    uint64_t r1[18], r4[18];
    uint8_t c1[48], c4[48];
    or_msm_pippenger(r1, p1.data(), sc.data(), n, 255, 0);
    or_msm_pippenger_mt(r4, p1.data(), sc.data(), n, 255, 0, 4);
    or_p1_compress(c1, r1);
    or_p1_compress(c4, r4);
    expect(std::memcmp(c1, c4, 48) == 0, "threaded msm", (int)n, 0);
    const size_t njobs = 6;
    std::vector<const uint64_t*> jp(njobs);
    std::vector<const uint8_t*> js(njobs);
    std::vector<size_t> jn(njobs), jb(njobs, 255);
    std::vector<uint64_t> rets(18 * njobs);
    for (size_t k = 0; k < njobs; k++) {
        jn[k] = 10 + 40 * k;
        jp[k] = p1.data();
        js[k] = sc.data();
    }
    or_msm_batch_mt(rets.data(), jp.data(), js.data(), jn.data(), jb.data(), njobs, 3);
    for (size_t k = 0; k < njobs; k++) {
        uint64_t r[18];
        uint8_t a[48], b[48];
        or_msm_pippenger(r, p1.data(), sc.data(), jn[k], 255, 0);
        or_p1_compress(a, r);
        or_p1_compress(b, &rets[18 * k]);
        expect(std::memcmp(a, b, 48) == 0, "threaded batch", (int)k, 0);
    }
}

}  // namespace

int main(int argc, char** argv) {
    const bool tsan = argc > 1 && std::strcmp(argv[1], "tsan") == 0;
    if (tsan) {
        check_threads();
    } else {
        check_host_products();
        check_g1_fold();
        check_g1_affine_and_inverse();
        check_g2_fold();
        check_divstep();
        check_fq_products();
        check_fq_bucket_sums();
    }
    std::printf("ok %d checks\n", g_checks);
    return 0;
}
