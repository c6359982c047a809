// Shared device definitions of the G1 Pippenger MSM stages on gfx950.
//
//   count       per tile (one window of one chunk of points): signed Booth
//               digits (WindowExtract / booth_encode,
//               blst/scalar/curve/multi_scalar.h:350-404, ec_mult.h:63-142)
//               counted into an LDS histogram of the window's buckets
//   prefix      per bucket: offsets of its chunks, and the bucket's total
//   scan        exclusive prefix sum of the totals -> bucket offsets
//   place       per tile again: LDS cursors put each (key, point | sign)
//               pair at its sorted position
//               (windows with more buckets than the LDS holds use a global
//               histogram with per-entry ranks instead: digits, scatter)
//   accumulate  fixed-length chunks of the sorted list per lane, XYZZ mixed
//               additions of gathered affine points (the bucket fill of
//               points_tile_pippenger, multi_scalar.cpp:806-865), G1 in the
//               28-bit limb field; buckets that cross chunk boundaries leave
//               partial sums in two slots/chunk
//   partials    buckets crossing one chunk boundary: one lane adds the two
//               partials; three or four pieces one lane, up to kWaveSpan one
//               wave, longer spans one workgroup per bucket (four-lane
//               additions in strides, then a tree), all in one launch
//   reduce      sum_b b * S_b per window (point_integrate_buckets,
//               multi_scalar.cpp:730-751): running sums over segments of L
//               buckets, then a tree of four-lane merges per window, G1 in
//               the 28-bit limb field
//
// Keys are (window * B + bucket) with B = 2^(c-1) buckets per window, so all
// windows go through one sort and one accumulate launch.
#pragma once
#include "ec.hip.h"

namespace bls {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Buckets whose accumulation left more than two pieces (they spanned more
// than one chunk boundary), listed by the accumulate kernel for the merge by
// piece count m: up to kLightSpan pieces one lane each, up to kWaveSpan one
// wave each, more one workgroup each.  counts: [wave, group, few] lengths,
// cleared before the accumulation.
constexpr uint32_t kLightSpan = 4;
constexpr uint32_t kWaveSpan = 128;

// A bucket of more than kWaveSpan pieces is merged by workgroups; when fewer
// such buckets are listed than the merge has workgroups for them, each
// bucket's pieces are cut into up to kHeavyParts parts (never more than its
// pieces), and parts[h] counts the finished parts of listed bucket h.
constexpr uint32_t kHeavyParts = 128;
static_assert(kHeavyParts <= kWaveSpan, "every part of a listed bucket holds a piece");

// list lengths (counts[0..2]), part counters, then the three lists
struct merge_lists_t {
    uint32_t* counts;
    uint32_t* parts;
    uint32_t* few;
    uint32_t* wave;
    uint32_t* group;
};
constexpr uint32_t kListHeadWords = 64 + kHeavyParts;

__device__ __forceinline__ void list_bucket(const merge_lists_t& L, uint32_t m, uint32_t key) {
    if (m <= kLightSpan) L.few[atomicAdd(&L.counts[2], 1u)] = key;
    else if (m <= kWaveSpan) L.wave[atomicAdd(&L.counts[0], 1u)] = key;
    else L.group[atomicAdd(&L.counts[1], 1u)] = key;
}


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

// G2 (Fp2 coordinates): the same layouts at twice the size
__device__ __forceinline__ void fp_from_uint4(fp& r, const uint4* v) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r.l[4 * k] = v[k].x; r.l[4 * k + 1] = v[k].y; r.l[4 * k + 2] = v[k].z; r.l[4 * k + 3] = v[k].w;
    }
}

__device__ __forceinline__ void load_aff(aff2& a, const uint4* __restrict__ pts, uint32_t idx) {
    const uint4* s = pts + (size_t)idx * 12;
    uint4 v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = s[k];
    fp_from_uint4(a.x.c0, v);
    fp_from_uint4(a.x.c1, v + 3);
    fp_from_uint4(a.y.c0, v + 6);
    fp_from_uint4(a.y.c1, v + 9);
}

__device__ __forceinline__ void store_xyzz(xyzz2* __restrict__ dst, const xyzz2& p) {
    uint4* d = reinterpret_cast<uint4*>(dst);
    const fp* f = &p.x.c0;
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            d[q * 3 + k] = make_uint4(f[q].l[4 * k], f[q].l[4 * k + 1], f[q].l[4 * k + 2], f[q].l[4 * k + 3]);
}

__device__ __forceinline__ void load_xyzz(xyzz2& p, const xyzz2* __restrict__ src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    fp* f = &p.x.c0;
#pragma unroll
    for (int q = 0; q < 8; q++) fp_from_uint4(f[q], s + 3 * q);
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

// len (<= 25) bits of a scalar starting at bit pos, read straight from its
// bytes; pos + len must not pass the scalar's nbits (so no bit at or above
// nbits is ever read, as with load_scalar's masking).  vec: 32-byte scalars
// on a 16-byte boundary, read as 32-bit words.
__device__ __forceinline__ uint32_t scalar_window(const uint8_t* __restrict__ s, bool vec, uint32_t pos,
                                                  uint32_t len) {
    if (vec) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(s);
        const uint32_t wi = pos >> 5, sh = pos & 31;
        uint64_t v = w[wi];
        if (sh + len > 32) v |= (uint64_t)w[wi + 1] << 32;
        return (uint32_t)(v >> sh) & ((1u << len) - 1u);
    }
    const uint32_t b0 = pos >> 3, b1 = (pos + len - 1) >> 3;
    uint32_t v = 0;
    for (uint32_t b = b0; b <= b1; b++) v |= (uint32_t)s[b] << (8 * (b - b0));
    return (v >> (pos & 7)) & ((1u << len) - 1u);
}

// booth_digit of the window at bit0 (<= nbits), reading only that window's
// bits of the scalar.
__device__ __forceinline__ int32_t booth_digit_at(const uint8_t* __restrict__ s, bool vec, uint32_t bit0,
                                                  uint32_t c, uint32_t nbits) {
    const uint32_t e = min(nbits - bit0 + 1, c);
    const uint32_t end = min(bit0 + c, nbits);
    uint32_t wval;
    if (bit0 == 0) wval = scalar_window(s, vec, 0, end) << 1;
    else wval = scalar_window(s, vec, bit0 - 1, end - bit0 + 1);
    const uint32_t sign = (wval >> e) & 1u;
    const int32_t val = (int32_t)((wval + 1) >> 1);
    return sign ? val - (int32_t)(1u << e) : val;
}

// ---------------------------------------------------------------- stages
//
// A launch covers a batch of MSMs ("jobs").  Job j owns the key range
// [kbase, kbase + nwin*B), the entry range [ebase, ebase + nwin*n) of the
// digit arrays, the points [pbase, pbase + n) of the flattened point
// enumeration and the groups (windows) [gbase, gbase + nwin).  A single MSM
// is a batch of one.

struct job_t {
    const uint8_t* scalars;   // device, n * nbytes
    const uint4* points;      // device, n affine points (96 bytes each)
    uint32_t n, nbytes, nbits, c, nwin, B, bit_base, vec;
    uint32_t kbase, ebase, pbase, gbase;
};

struct group_t {
    uint32_t kb;   // first key (bucket 0 of this window)
    uint32_t B;    // buckets
    uint32_t L;    // buckets per reduction segment
    uint32_t sb;   // first segment index in the H array
    uint32_t P;    // segments
    uint32_t pad[3];
};

// Sort tiles of one group (a window of a job): the job's points are cut into
// nch chunks of ch points; chunk k's bucket counts live at
// counts[cb + k * B + bucket].
struct sgroup_t {
    uint32_t job, w;
    uint32_t cb;
    uint32_t nch, ch;
    uint32_t pad[3];
};

constexpr uint32_t kSortThreads = 512;
// points per thread and step of the count and place loops
constexpr uint32_t kSortBatch = 4;

// Two-pass ("binned") sort for batches whose windows have 2^6 .. 2^15
// buckets, mostly 2^12 or more: bins of 2^kBinBits consecutive buckets, tiles
// of kBinTile points (kBinTile / kSortThreads digits per thread)
constexpr uint32_t kBinBits = 6;
constexpr uint32_t kBinTile = 8192;
constexpr uint32_t kBinMaxBins = kSortThreads;

// nodes one block of the reduction tree merges into one
constexpr uint32_t kTreeNodes = 64;

// largest j with base[j] <= t (base ascending, base[0] = 0)
__device__ __forceinline__ uint32_t find_owner(const uint32_t* __restrict__ base, uint32_t cnt, uint32_t t) {
    uint32_t lo = 0, hi = cnt;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// exclusive scan, 1024 threads x 4 elements per block
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

}  // namespace bls
