// Device stages of the G1 Pippenger MSM on gfx950.
//
//   digits      signed Booth digits of every (window, scalar), a global
//               histogram of (window, bucket) keys with the rank of each
//               entry inside its key (WindowExtract / booth_encode,
//               blst/scalar/curve/multi_scalar.h:350-404, ec_mult.h:63-142)
//   scan        exclusive prefix sum of the histogram -> bucket offsets
//   scatter     counting-sort placement of (key, point | sign) pairs
//   accumulate  fixed-length chunks of the sorted list per lane, XYZZ mixed
//               additions of gathered affine points (the bucket fill of
//               points_tile_pippenger, multi_scalar.cpp:806-865); buckets that
//               cross chunk boundaries leave partial sums in two slots/chunk
//   partials    log-depth pairwise tree over the partial sums of a bucket
//   reduce      sum_b b * S_b per window (point_integrate_buckets,
//               multi_scalar.cpp:730-751) as segment running sums plus a
//               per-segment scalar weight, then a pairwise tree per window
//
// Keys are (window * B + bucket) with B = 2^(c-1) buckets per window, so all
// windows go through one sort and one accumulate launch.
#pragma once
#include "ec.hip.h"

namespace bls {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

struct win_params {
    uint32_t n;         // points
    uint32_t nbytes;    // bytes per scalar
    uint32_t nbits;     // meaningful scalar bits
    uint32_t c;         // window bits
    uint32_t nwin;      // windows
    uint32_t bit_base;  // bit offset of window 0
    uint32_t B;         // buckets per window
    uint32_t vec_scalars;  // scalars are 32 bytes each and 16-byte aligned
};

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ void load_aff(aff& a, const uint4* __restrict__ pts, uint32_t idx) {
    const uint4* s = pts + (size_t)idx * 6;
    uint4 v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = s[k];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a.x.l[4 * k + 0] = v[k].x; a.x.l[4 * k + 1] = v[k].y;
        a.x.l[4 * k + 2] = v[k].z; a.x.l[4 * k + 3] = v[k].w;
        a.y.l[4 * k + 0] = v[k + 3].x; a.y.l[4 * k + 1] = v[k + 3].y;
        a.y.l[4 * k + 2] = v[k + 3].z; a.y.l[4 * k + 3] = v[k + 3].w;
    }
}

__device__ __forceinline__ void store_xyzz(xyzz* __restrict__ dst, const xyzz& p) {
    uint4* d = reinterpret_cast<uint4*>(dst);
    const fp* f = &p.x;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            d[q * 3 + k] = make_uint4(f[q].l[4 * k], f[q].l[4 * k + 1], f[q].l[4 * k + 2], f[q].l[4 * k + 3]);
}

__device__ __forceinline__ void load_xyzz(xyzz& p, const xyzz* __restrict__ src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    fp* f = &p.x;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            uint4 v = s[q * 3 + k];
            f[q].l[4 * k] = v.x; f[q].l[4 * k + 1] = v.y; f[q].l[4 * k + 2] = v.z; f[q].l[4 * k + 3] = v.w;
        }
}

// Scalar i as nine 32-bit words (the ninth is zero padding), bits at and
// above nbits cleared: the reference reads only ceil(nbits/8) bytes and
// masks each window to nbits (WindowExtract::initialize, window_end).
__device__ __forceinline__ void load_scalar(uint32_t k[9], const uint8_t* __restrict__ s,
                                            uint32_t nbytes, uint32_t nbits, bool vec) {
#pragma unroll
    for (int w = 0; w < 9; w++) k[w] = 0;
    if (vec) {
        const uint4* s4 = reinterpret_cast<const uint4*>(s);
        uint4 a = s4[0], b = s4[1];
        k[0] = a.x; k[1] = a.y; k[2] = a.z; k[3] = a.w;
        k[4] = b.x; k[5] = b.y; k[6] = b.z; k[7] = b.w;
    } else {
        for (uint32_t i = 0; i < nbytes; i++) k[i >> 2] |= (uint32_t)s[i] << (8 * (i & 3));
    }
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const int lo = 32 * w;
        if ((int)nbits <= lo) k[w] = 0;
        else if ((int)nbits < lo + 32) k[w] &= (1u << (nbits - lo)) - 1u;
    }
}

// len (<= 25) bits of k starting at bit pos (pos + len <= 257)
__device__ __forceinline__ uint32_t scalar_bits(const uint32_t k[9], uint32_t pos, uint32_t len) {
    const uint32_t w = pos >> 5, sh = pos & 31;
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        lo = (j == (int)w) ? k[j] : lo;
        hi = (j == (int)w + 1) ? k[j] : hi;
    }
    const uint64_t v = (((uint64_t)hi << 32) | lo) >> sh;
    return (uint32_t)v & ((len >= 32) ? 0xffffffffu : ((1u << len) - 1u));
}

// Signed Booth digit of the window starting at bit0 (< nbits) of width c:
// bits [bit0-1, min(bit0+c, nbits)) with a zero below bit 0, encoded with
// e = min(nbits - bit0 + 1, c) as in booth_encode: digit in [-2^(e-1), 2^(e-1)].
__device__ __forceinline__ int32_t booth_digit(const uint32_t k[9], uint32_t bit0, uint32_t c,
                                               uint32_t nbits) {
    const uint32_t e = min(nbits - bit0 + 1, c);
    const uint32_t end = min(bit0 + c, nbits);
    uint32_t wval;
    if (bit0 == 0) wval = scalar_bits(k, 0, end) << 1;
    else wval = scalar_bits(k, bit0 - 1, end - bit0 + 1);
    const uint32_t sign = (wval >> e) & 1u;
    const int32_t val = (int32_t)((wval + 1) >> 1);
    return sign ? val - (int32_t)(1u << e) : val;
}

// ---------------------------------------------------------------- stages

__global__ void __launch_bounds__(256)
k_digits(win_params P, const uint8_t* __restrict__ scalars,
                         const uint4* __restrict__ points, uint32_t* __restrict__ ent,
                         uint32_t* __restrict__ rank, uint32_t* __restrict__ hist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    uint32_t k[9];
    load_scalar(k, scalars + (size_t)i * P.nbytes, P.nbytes, P.nbits, P.vec_scalars);
    // an all-zero affine point is infinity and contributes nothing
    // (Pointxyzz::dadd with an all-zero affine input, point.cpp:2036-2039)
    const uint4* pp = points + (size_t)i * 6;
    uint32_t nz = 0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        uint4 v = pp[q];
        nz |= v.x | v.y | v.z | v.w;
    }
    for (uint32_t w = 0; w < P.nwin; w++) {
        const uint32_t bit0 = P.bit_base + w * P.c;
        // bit0 == nbits is the top window of an nbits that c divides: it holds
        // only the carry of the window below (WindowExtract asserts
        // window_idx <= nbits / window, multi_scalar.h:365)
        int32_t d = (nz && bit0 <= P.nbits) ? booth_digit(k, bit0, P.c, P.nbits) : 0;
        const size_t slot = (size_t)w * P.n + i;
        if (d != 0) {
            const uint32_t neg = d < 0;
            const uint32_t b = (uint32_t)(neg ? -d : d) - 1u;
            const uint32_t key = w * P.B + b;
            rank[slot] = atomicAdd(&hist[key], 1u);
            ent[slot] = key | (neg << 31);
        } else {
            ent[slot] = kInvalid;
        }
    }
}

__global__ void k_digits_debug(const uint8_t* __restrict__ scalars, uint32_t n, uint32_t nbytes,
                               uint32_t nbits, uint32_t bit0, uint32_t c, int32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k[9];
    load_scalar(k, scalars + (size_t)i * nbytes, nbytes, nbits, false);
    out[i] = booth_digit(k, bit0, c, nbits);
}

// exclusive scan, 1024 threads x 4 elements per block
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
             uint32_t* __restrict__ tile_sums, uint32_t N) {
    __shared__ uint32_t sh[kScanThreads];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems], run = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        v[j] = (base + j < N) ? in[base + j] : 0u;
        run += v[j];
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        uint32_t t = (threadIdx.x >= (uint32_t)off) ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t excl = sh[threadIdx.x] - run;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        if (base + j < N) out[base + j] = excl;
        excl += v[j];
    }
    if (threadIdx.x == kScanThreads - 1 && tile_sums) tile_sums[blockIdx.x] = sh[threadIdx.x];
}

__global__ void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_offs, uint32_t N) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out[i] += tile_offs[i / kScanTile];
}

__global__ void __launch_bounds__(256)
k_scatter(win_params P, const uint32_t* __restrict__ ent, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ off, uint32_t* __restrict__ skey,
                          uint32_t* __restrict__ sval) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)P.nwin * P.n) return;
    const uint32_t e = ent[t];
    if (e == kInvalid) return;
    const uint32_t key = e & 0x7fffffffu;
    const uint32_t i = (uint32_t)(t % P.n);
    const uint32_t pos = off[key] + rank[t];
    skey[pos] = key;
    sval[pos] = i | (e & 0x80000000u);
}

// One lane per chunk of T sorted entries.  A bucket segment that starts and
// ends inside the chunk is final and goes to buckets[key].  Otherwise, with
// k0 the chunk holding the bucket's first entry: the segment in chunk k0 goes
// to slot 2*k0+1, the segment in any later chunk k to slot 2*k.
__global__ void __launch_bounds__(256)
k_accumulate(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval,
             const uint32_t* __restrict__ total, const uint4* __restrict__ points,
             xyzz* __restrict__ buckets, uint32_t* __restrict__ slot_key,
             xyzz* __restrict__ slots, uint32_t T, uint32_t nchunks) {
    const uint32_t chunk = blockIdx.x * blockDim.x + threadIdx.x;
    if (chunk >= nchunks) return;
    slot_key[2 * chunk] = kInvalid;
    slot_key[2 * chunk + 1] = kInvalid;
    const uint32_t M = *total;
    const uint32_t start = chunk * T;
    if (start >= M) return;
    const uint32_t end = min(start + T, M);
    uint32_t cur = skey[start];
    const bool before = start > 0 && skey[start - 1] == cur;
    bool first = true;
    xyzz acc;
    xyzz_set_inf(acc);
    for (uint32_t e = start; e < end; e++) {
        const uint32_t key = skey[e];
        if (key != cur) {
            if (first && before) {
                slot_key[2 * chunk] = cur;
                store_xyzz(&slots[2 * chunk], acc);
            } else {
                store_xyzz(&buckets[cur], acc);
            }
            first = false;
            cur = key;
            xyzz_set_inf(acc);
        }
        const uint32_t v = sval[e];
        aff a;
        load_aff(a, points, v & 0x7fffffffu);
        if (v >> 31) fp_neg(a.y, a.y);
        xyzz_madd(acc, a);
    }
    const bool after = end < M && skey[end] == cur;
    if (first && before) {
        slot_key[2 * chunk] = cur;
        store_xyzz(&slots[2 * chunk], acc);
    } else if (after) {
        slot_key[2 * chunk + 1] = cur;
        store_xyzz(&slots[2 * chunk + 1], acc);
    } else {
        store_xyzz(&buckets[cur], acc);
    }
}

// Round with distance d of the pairwise tree over the partials of one bucket:
// partial j (j = 0 in slot 2*k0+1, j >= 1 in slot 2*(k0+j)) absorbs partial
// j+d when j is a multiple of 2d.
__global__ void __launch_bounds__(256)
k_partial_round(const uint32_t* __restrict__ slot_key, xyzz* __restrict__ slots,
                                const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                                uint32_t T, uint32_t nslots, uint32_t d) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint32_t key = slot_key[s];
    if (key == kInvalid) return;
    const uint32_t k0 = off[key] / T;
    const uint32_t k1 = (off[key] + cnt[key] - 1) / T;
    const uint32_t j = (s & 1) ? 0u : (s >> 1) - k0;
    if (j % (2 * d)) return;
    const uint32_t jj = j + d;
    if (jj > k1 - k0) return;
    xyzz a, b;
    load_xyzz(a, &slots[s]);
    load_xyzz(b, &slots[2 * (k0 + jj)]);
    xyzz_add(a, a, b);
    store_xyzz(&slots[s], a);
}

__global__ void __launch_bounds__(256)
k_partial_final(const uint32_t* __restrict__ slot_key, const xyzz* __restrict__ slots,
                                xyzz* __restrict__ buckets, uint32_t nslots) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots || !(s & 1)) return;
    const uint32_t key = slot_key[s];
    if (key == kInvalid) return;
    xyzz a;
    load_xyzz(a, &slots[s]);
    store_xyzz(&buckets[key], a);
}

// Segment s of L buckets of window w (bucket b holds digit b+1):
// H = sum_t (t+1) S_(sL+t) + (sL) * sum_t S_(sL+t), i.e. the segment's share
// of sum_b (b+1) S_b.
__global__ void __launch_bounds__(256)
k_reduce_segments(const xyzz* __restrict__ buckets, xyzz* __restrict__ H, uint32_t B, uint32_t L,
                  uint32_t nwin) {
    const uint32_t P = B / L;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwin * P) return;
    const uint32_t w = t / P, s = t % P;
    const size_t base = (size_t)w * B + (size_t)s * L;
    xyzz acc, tot, b;
    xyzz_set_inf(acc);
    xyzz_set_inf(tot);
    for (int j = (int)L - 1; j >= 0; j--) {
        load_xyzz(b, &buckets[base + j]);
        xyzz_add(acc, acc, b);
        xyzz_add(tot, tot, acc);
    }
    const uint32_t m = s * L;
    if (m) {
        xyzz r;
        xyzz_set_inf(r);
        for (int bit = 31 - __builtin_clz(m); bit >= 0; bit--) {
            xyzz_dbl_inl(r, r);
            if ((m >> bit) & 1u) xyzz_add(r, r, acc);
        }
        xyzz_add(tot, tot, r);
    }
    store_xyzz(&H[t], tot);
}

// H[w*P + i*2*step] += H[w*P + i*2*step + step]
__global__ void __launch_bounds__(256)
k_pair_sum(xyzz* __restrict__ H, uint32_t P, uint32_t nwin, uint32_t step) {
    const uint32_t per = P / (2 * step);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwin * per) return;
    const uint32_t w = t / per, i = t % per;
    const size_t a = (size_t)w * P + (size_t)i * 2 * step;
    xyzz x, y;
    load_xyzz(x, &H[a]);
    load_xyzz(y, &H[a + step]);
    xyzz_add(x, x, y);
    store_xyzz(&H[a], x);
}

// ------------------------------------------------ fixed-base and affine

// [k_i]G from a comb table T[m][j-1] = [j * 2^(8m)] G (m < nbytes, j < 256):
// one mixed addition per nonzero scalar byte.
__global__ void __launch_bounds__(256)
k_fixed_base(const uint8_t* __restrict__ scalars, uint32_t n, uint32_t nbytes, uint32_t nbits,
             const uint4* __restrict__ table, xyzz* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k[9];
    load_scalar(k, scalars + (size_t)i * nbytes, nbytes, nbits, false);
    xyzz acc;
    xyzz_set_inf(acc);
    for (uint32_t m = 0; m < nbytes; m++) {
        const uint32_t b = scalar_bits(k, 8 * m, 8);
        if (b) {
            aff a;
            load_aff(a, table, m * 255 + (b - 1));
            xyzz_madd(acc, a);
        }
    }
    store_xyzz(&out[i], acc);
}

// XYZZ -> affine, Q points per lane sharing one inversion of the product of
// ZZ*ZZZ (= Z^5); prefix products are parked in the output's x field.
__global__ void __launch_bounds__(256)
k_to_affine(const xyzz* __restrict__ in, aff* __restrict__ out, uint32_t n, uint32_t Q) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t begin = lane * Q;
    if (begin >= n) return;
    const uint32_t end = min(begin + Q, n);
    fp acc, t;
    fp_one(acc);
    for (uint32_t i = begin; i < end; i++) {
        xyzz p;
        load_xyzz(p, &in[i]);
        out[i].x = acc;
        if (!xyzz_is_inf(p)) {
            fp_mul(t, p.zz, p.zzz);
            fp_mul(acc, acc, t);
        }
    }
    fp ia;
    fp_inv(ia, acc);
    for (uint32_t i = end; i-- > begin;) {
        xyzz p;
        load_xyzz(p, &in[i]);
        aff o;
        if (xyzz_is_inf(p)) {
            fp_zero(o.x);
            fp_zero(o.y);
        } else {
            fp it, izz, izzz;
            const fp pre = out[i].x;
            fp_mul(t, p.zz, p.zzz);
            fp_mul(it, ia, pre);
            fp_mul(ia, ia, t);
            fp_mul(izz, it, p.zzz);
            fp_mul(izzz, it, p.zz);
            fp_mul(o.x, p.x, izz);
            fp_mul(o.y, p.y, izzz);
        }
        out[i] = o;
    }
}

}  // namespace bls
