// Host build of the G1 accumulation's 28-bit limb field
// (temp-cat-blst_amd/csrc/fq.hip.h, compiled here by g++ as plain C++) for
// tests/test_fq.py: the operations of blst_hip_fq_ops by the same op numbers,
// and a bucket sum over a list of affine points as the accumulation kernel
// forms it.
#include "../temp-cat-blst_amd/csrc/fq.hip.h"
#include <string.h>

using namespace bls;

extern "C" {

void fqh_ops(int op, uint32_t* out, const uint32_t* A, const uint32_t* B, const uint32_t* C, const uint32_t* D,
             size_t n) {
    for (size_t i = 0; i < n; i++) {
        fq a, b, c, d, r;
        memcpy(a.l, A + 14 * i, 56);
        memcpy(b.l, B + 14 * i, 56);
        memcpy(c.l, C + 14 * i, 56);
        memcpy(d.l, D + 14 * i, 56);
        fq_zero(r);
        switch (op) {
        case 0: fq_mul(r, a, b); break;
        case 1: fq_sqr(r, a); break;
        case 2: fq_mul_sum(r, a, b, c, d); break;
        case 3: fq_reduce(r, a); break;
        case 4: fq_sub_lazy(r, a, b); break;
        case 5: fq_neg_lazy(r, a); break;
        case 6: fq_x3(r, a, b, c, (d.l[0] & 1) != 0); break;
        case 7: r.l[0] = fq_is_zero(a) ? 1u : 0u; break;
        case 8: fq_from_words<8>(r, a.l); break;
        case 9: fq_from_words<4>(r, a.l); break;
        case 10: fq_from_words<0>(r, a.l); break;
        case 11: fq_to_words(r.l, a); break;
        case 12: fq_times4(r, a); break;
        case 13: fq_reduce_wide(r, a); break;
        default: fq_canon(r, a); break;
        }
        memcpy(out + 14 * i, r.l, 56);
    }
}

// io: the 32-bit path XYZZ word image of the starting sum (48 words, all zero
// for infinity), replaced by the sum's; pts: n affine points of 24 words
// (canonical Montgomery words, all zero = infinity), neg[k] negates point k
void fqh_accumulate(uint32_t* io, const uint32_t* pts, const uint8_t* neg, size_t n) {
    static const uint32_t P2[12] = {0xffff5556u, 0x73fdffffu, 0x62a7ffffu, 0x3d57fffdu, 0xed61ec48u, 0xce61a541u,
                                    0xe70a257eu, 0xc8ee9709u, 0x869759aeu, 0x96374f6cu, 0x72ffcd34u, 0x340223d4u};
    fq_xyzz acc;
    fq_xyzz_from_words(acc, io, io + 12, io + 24, io + 36);
    for (size_t k = 0; k < n; k++) {
        uint32_t x[12], y[12], z = 0;
        memcpy(x, pts + 24 * k, 48);
        memcpy(y, pts + 24 * k + 12, 48);
        for (int i = 0; i < 12; i++) z |= x[i] | y[i];
        if (!z) continue;
        if (neg[k]) {   // 2p - y, as the kernel negates
            uint64_t borrow = 0;
            for (int i = 0; i < 12; i++) {
                const uint64_t t = (uint64_t)P2[i] - y[i] - borrow;
                y[i] = (uint32_t)t;
                borrow = (t >> 63) & 1;
            }
        }
        fq fx, fy16;
        fq_from_words<8>(fx, x);
        fq_from_words<4>(fy16, y);
        fq_xyzz_madd(acc, fx, fy16);
    }
    fq_xyzz_to_words(io, io + 12, io + 24, io + 36, acc);
}

}
