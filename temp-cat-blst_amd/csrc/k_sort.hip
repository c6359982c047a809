// Digit, histogram, scan and scatter stages: the device counting sort of
// (bucket key, point | sign) pairs.
#include "msm_common.hip.h"
#include "launch.h"

#include <algorithm>

namespace bls {

__global__ void __launch_bounds__(256)
k_digits(const job_t* __restrict__ jobs, const uint32_t* __restrict__ pbase, uint32_t njobs,
         uint32_t total_points, uint32_t* __restrict__ ent, uint32_t* __restrict__ rank,
         uint32_t* __restrict__ hist) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total_points) return;
    const uint32_t j = find_owner(pbase, njobs, t);
    const job_t J = jobs[j];
    const uint32_t i = t - J.pbase;
    uint32_t k[9];
    load_scalar(k, J.scalars + (size_t)i * J.nbytes, J.nbytes, J.nbits, J.vec);
    // an all-zero affine point is infinity and contributes nothing
    // (Pointxyzz::dadd with an all-zero affine input, point.cpp:2036-2039)
    const uint4* pp = J.points + (size_t)i * 6;
    uint32_t nz = 0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        uint4 v = pp[q];
        nz |= v.x | v.y | v.z | v.w;
    }
    for (uint32_t w = 0; w < J.nwin; w++) {
        const uint32_t bit0 = J.bit_base + w * J.c;
        // bit0 == nbits is the top window of an nbits that c divides: it holds
        // only the carry of the window below (WindowExtract asserts
        // window_idx <= nbits / window, multi_scalar.h:365)
        const int32_t d = (nz && bit0 <= J.nbits) ? booth_digit(k, bit0, J.c, J.nbits) : 0;
        const size_t slot = (size_t)J.ebase + (size_t)w * J.n + i;
        if (d != 0) {
            const uint32_t neg = d < 0;
            const uint32_t b = (uint32_t)(neg ? -d : d) - 1u;
            const uint32_t key = J.kbase + w * J.B + b;
            rank[slot] = atomicAdd(&hist[key], 1u);
            ent[slot] = key | (neg << 31);
        } else {
            ent[slot] = kInvalid;
        }
    }
}

// Tile = chunk k of the points of group g, from the tile table (one launch
// per LDS size class).  LDS holds the B bucket counters of the group's
// window.
__global__ void __launch_bounds__(kSortThreads)
k_count(const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles,
        uint32_t* __restrict__ counts) {
    extern __shared__ uint32_t bins[];
    const uint2 tk = tiles[blockIdx.x];
    const uint32_t g = tk.x, k = tk.y;
    const sgroup_t S = sg[g];
    const job_t& J = jobs[S.job];
    const uint32_t B = J.B, c = J.c, nbits = J.nbits, nbytes = J.nbytes;
    const bool vec = J.vec;
    const uint8_t* scalars = J.scalars;
    for (uint32_t b = threadIdx.x; b < B; b += blockDim.x) bins[b] = 0;
    __syncthreads();
    const uint32_t i0 = k * S.ch, i1 = min(i0 + S.ch, J.n);
    const uint32_t bit0 = J.bit_base + S.w * c;
    if (bit0 <= nbits) {
        // kSortBatch digits per step: their scalar loads are issued together
        for (uint32_t i = i0 + threadIdx.x; i < i1; i += kSortBatch * blockDim.x) {
            int32_t d[kSortBatch];
#pragma unroll
            for (uint32_t q = 0; q < kSortBatch; q++) {
                const uint32_t iq = i + q * blockDim.x;
                d[q] = iq < i1 ? booth_digit_at(scalars + (size_t)iq * nbytes, vec, bit0, c, nbits) : 0;
            }
#pragma unroll
            for (uint32_t q = 0; q < kSortBatch; q++)
                if (d[q]) atomicAdd(&bins[(uint32_t)(d[q] < 0 ? -d[q] : d[q]) - 1u], 1u);
        }
    }
    __syncthreads();
    uint32_t* out = counts + S.cb + (size_t)k * B;
    for (uint32_t b = threadIdx.x; b < B; b += blockDim.x) out[b] = bins[b];
}

// Per key: chunk counts -> exclusive running offsets inside the key, and the
// key's total into hist (hist[nkeys] = 0 closes the scan).
__global__ void __launch_bounds__(256)
k_prefix(const group_t* __restrict__ groups, const sgroup_t* __restrict__ sg, const uint32_t* __restrict__ gkb,
         uint32_t ngroups, uint32_t nkeys, uint32_t* __restrict__ counts, uint32_t* __restrict__ hist) {
    const uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key > nkeys) return;
    if (key == nkeys) {
        hist[key] = 0;
        return;
    }
    const uint32_t g = find_owner(gkb, ngroups, key);
    const uint32_t B = groups[g].B;
    const sgroup_t S = sg[g];
    uint32_t* p = counts + S.cb + (key - groups[g].kb);
    uint32_t run = 0;
    for (uint32_t k = 0; k < S.nch; k++, p += B) {
        const uint32_t v = *p;
        *p = run;
        run += v;
    }
    hist[key] = run;
}

// Per tile again: LDS cursors start at the bucket offset plus the chunk's
// offset inside the bucket and hand out positions.
__global__ void __launch_bounds__(kSortThreads)
k_place(const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles,
        const uint32_t* __restrict__ counts, const uint32_t* __restrict__ off, uint2* __restrict__ sorted) {
    extern __shared__ uint32_t cursor[];
    const uint2 tk = tiles[blockIdx.x];
    const uint32_t g = tk.x, k = tk.y;
    const sgroup_t S = sg[g];
    const job_t& J = jobs[S.job];
    const uint32_t B = J.B, c = J.c, nbits = J.nbits, nbytes = J.nbytes;
    const bool vec = J.vec;
    const uint8_t* scalars = J.scalars;
    const uint32_t kb = J.kbase + S.w * B;
    const uint32_t* cnt = counts + S.cb + (size_t)k * B;
    for (uint32_t b = threadIdx.x; b < B; b += blockDim.x) cursor[b] = off[kb + b] + cnt[b];
    __syncthreads();
    const uint32_t i0 = k * S.ch, i1 = min(i0 + S.ch, J.n);
    const uint32_t bit0 = J.bit_base + S.w * c;
    if (bit0 > nbits) return;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += kSortBatch * blockDim.x) {
        int32_t d[kSortBatch];
#pragma unroll
        for (uint32_t q = 0; q < kSortBatch; q++) {
            const uint32_t iq = i + q * blockDim.x;
            d[q] = iq < i1 ? booth_digit_at(scalars + (size_t)iq * nbytes, vec, bit0, c, nbits) : 0;
        }
#pragma unroll
        for (uint32_t q = 0; q < kSortBatch; q++) {
            if (d[q]) {
                const uint32_t neg = d[q] < 0;
                const uint32_t b = (uint32_t)(neg ? -d[q] : d[q]) - 1u;
                const uint32_t pos = atomicAdd(&cursor[b], 1u);
                sorted[pos] = make_uint2(kb + b, (i + q * blockDim.x) | (neg << 31));
            }
        }
    }
}

// ---------------------------------------------------------------- binned sort
//
// For windows of 2^12 .. 2^15 buckets a tile of the one-pass sort puts about
// two entries into each bucket's run, so k_place's 8-byte stores land one or
// two per cache line.  The binned sort groups 2^kBinBits buckets into a bin:
//   k_bin_count  per tile (window x kBinTile points): bin histogram in LDS,
//                stored bin-major, counts[cb + bin * nch + tile]
//   scan         one exclusive scan of all counts: the start of every
//                (bin, tile) run in the sorted order, bins in key order
//   k_bin_place  per tile: entries staged in LDS in bin order, then each
//                bin's run written contiguously into tmp
//   k_bin_fine   per bin: its buckets counted in LDS (the per-bucket count
//                and offset tables of the later stages) and its entries
//                placed in bucket order inside the bin's region; a bin of up
//                to 4096 entries is read from HBM once, into LDS
// Keys of a binned layout start every window at a multiple of 2^kBinBits,
// so the global bin of a key is key >> kBinBits.

__device__ __forceinline__ uint32_t bin_tile_digits(const job_t& J, const sgroup_t& S, uint32_t k,
                                                    int32_t d[kBinTile / kSortThreads]) {
    const uint32_t i0 = k * kBinTile + threadIdx.x, i1 = J.n;
    const uint32_t bit0 = J.bit_base + S.w * J.c;
#pragma unroll
    for (uint32_t q = 0; q < kBinTile / kSortThreads; q++) {
        const uint32_t i = i0 + q * kSortThreads;
        d[q] = (i < i1 && bit0 <= J.nbits)
                   ? booth_digit_at(J.scalars + (size_t)i * J.nbytes, J.vec, bit0, J.c, J.nbits)
                   : 0;
    }
    return i0;
}

__global__ void __launch_bounds__(kSortThreads)
k_bin_count(const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles,
            uint32_t* __restrict__ counts, uint32_t ncounts) {
    __shared__ uint32_t bins[kBinMaxBins];
    const uint2 tk = tiles[blockIdx.x];
    const sgroup_t S = sg[tk.x];
    const job_t& J = jobs[S.job];
    const uint32_t nbins = J.B >> kBinBits;
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) bins[b] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[ncounts] = 0;   // closes the scan
    __syncthreads();
    int32_t d[kBinTile / kSortThreads];
    bin_tile_digits(J, S, tk.y, d);
#pragma unroll
    for (uint32_t q = 0; q < kBinTile / kSortThreads; q++)
        if (d[q]) atomicAdd(&bins[((uint32_t)(d[q] < 0 ? -d[q] : d[q]) - 1u) >> kBinBits], 1u);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) counts[S.cb + (size_t)b * S.nch + tk.y] = bins[b];
}

// exclusive scan of v over the block (blockDim.x == kSortThreads), LDS scratch
// of kSortThreads / 64 words; returns the exclusive prefix, *total the sum
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < kSortThreads / 64; w++) {
        const uint32_t s = wsum[w];
        before += w < wave ? s : 0u;
        all += s;
    }
    *total = all;
    return before + x - v;
}

__global__ void __launch_bounds__(kSortThreads)
k_bin_place(const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles,
            const uint32_t* __restrict__ pos, uint2* __restrict__ tmp) {
    __shared__ uint2 stage[kBinTile];
    __shared__ uint32_t start[kBinMaxBins], cursor[kBinMaxBins], wsum[kSortThreads / 64];
    const uint2 tk = tiles[blockIdx.x];
    const sgroup_t S = sg[tk.x];
    const job_t& J = jobs[S.job];
    const uint32_t nbins = J.B >> kBinBits;
    const uint32_t kb = J.kbase + S.w * J.B;
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) cursor[b] = 0;
    __syncthreads();
    int32_t d[kBinTile / kSortThreads];
    const uint32_t i0 = bin_tile_digits(J, S, tk.y, d);
#pragma unroll
    for (uint32_t q = 0; q < kBinTile / kSortThreads; q++)
        if (d[q]) atomicAdd(&cursor[((uint32_t)(d[q] < 0 ? -d[q] : d[q]) - 1u) >> kBinBits], 1u);
    __syncthreads();
    // nbins <= kSortThreads: one bin per thread
    uint32_t m;
    const uint32_t c = threadIdx.x < nbins ? cursor[threadIdx.x] : 0u;
    const uint32_t s0 = block_scan(c, wsum, &m);
    if (threadIdx.x < nbins) {
        start[threadIdx.x] = s0;
        cursor[threadIdx.x] = s0;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kBinTile / kSortThreads; q++) {
        if (d[q]) {
            const uint32_t neg = d[q] < 0;
            const uint32_t b = (uint32_t)(neg ? -d[q] : d[q]) - 1u;
            const uint32_t at = atomicAdd(&cursor[b >> kBinBits], 1u);
            stage[at] = make_uint2(kb + b, (i0 + q * kSortThreads) | (neg << 31));
        }
    }
    __syncthreads();
    // bin j's entries are stage[start[j] .. start[j] + count): consecutive
    // slots go to consecutive positions of the bin's run
    for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
        const uint2 v = stage[e];
        const uint32_t j = (v.x - kb) >> kBinBits;
        tmp[pos[S.cb + (size_t)j * S.nch + tk.y] + (e - start[j])] = v;
    }
}

// one workgroup per bin; gkb[g] is the first key of group g
__global__ void __launch_bounds__(256)
k_bin_fine(const sgroup_t* __restrict__ sg, const uint32_t* __restrict__ gkb, uint32_t ngroups,
           const uint32_t* __restrict__ pos, const uint2* __restrict__ tmp, uint2* __restrict__ sorted,
           uint32_t* __restrict__ cnt, uint32_t* __restrict__ off, uint32_t nkeys, uint32_t ncounts) {
    constexpr uint32_t F = 1u << kBinBits;
    // a bin of up to kFineStage entries (about twice the average at 2^20) is
    // read from HBM once, into LDS, and counted and placed from there
    constexpr uint32_t kFineStage = 4096;
    __shared__ uint32_t fcount[F], cursor[F];
    __shared__ uint2 stage[kFineStage];
    const uint32_t key0 = blockIdx.x << kBinBits;
    const uint32_t g = find_owner(gkb, ngroups, key0);
    const sgroup_t S = sg[g];
    const uint32_t j = (key0 - gkb[g]) >> kBinBits;
    const uint32_t lo = pos[S.cb + (size_t)j * S.nch], hi = pos[S.cb + (size_t)(j + 1) * S.nch];
    if (threadIdx.x < F) fcount[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) off[nkeys] = pos[ncounts];   // total entries
    __syncthreads();
    const bool staged = hi - lo <= kFineStage;   // uniform over the block
    if (staged) {
        for (uint32_t e = threadIdx.x; e < hi - lo; e += blockDim.x) {
            const uint2 v = tmp[lo + e];
            stage[e] = v;
            atomicAdd(&fcount[v.x & (F - 1)], 1u);
        }
    } else {
        for (uint32_t e = lo + threadIdx.x; e < hi; e += blockDim.x) atomicAdd(&fcount[tmp[e].x & (F - 1)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // F == 64: one wave scans the bin's bucket counts
        const uint32_t v = fcount[threadIdx.x];
        uint32_t x = v;
#pragma unroll
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (threadIdx.x >= o) x += y;
        }
        cursor[threadIdx.x] = x - v;
        cnt[key0 + threadIdx.x] = v;
        off[key0 + threadIdx.x] = lo + x - v;
    }
    __syncthreads();
    if (staged) {
        for (uint32_t e = threadIdx.x; e < hi - lo; e += blockDim.x) {
            const uint2 v = stage[e];
            sorted[lo + atomicAdd(&cursor[v.x & (F - 1)], 1u)] = v;
        }
    } else {
        for (uint32_t e = lo + threadIdx.x; e < hi; e += blockDim.x) {
            const uint2 v = tmp[e];
            sorted[lo + atomicAdd(&cursor[v.x & (F - 1)], 1u)] = v;
        }
    }
}
static_assert((1u << kBinBits) == 64, "k_bin_fine scans one bin with one wave");

// Montgomery-form Fr scalars (blst_fr: x = k * 2^256 mod r, four 64-bit
// little-endian limbs) to canonical 32-byte little-endian scalars k =
// x * 2^-256 mod r, in place: eight rounds of word-wise Montgomery reduction
// and one conditional subtraction (Fr::to_scaler, blst/scalar/fields/Fr.h:445).
__global__ void __launch_bounds__(256)
k_fr_to_scalar(uint32_t* __restrict__ io, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr uint32_t R[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                               0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
    constexpr uint32_t N0 = 0xffffffffu;   // -r^-1 mod 2^32
    uint32_t* w = io + 8 * (size_t)i;
    uint32_t t[9];
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = w[j];
    t[8] = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t m = t[0] * N0;
        uint64_t c = ((uint64_t)m * R[0] + t[0]) >> 32;
#pragma unroll
        for (int j = 1; j < 8; j++) {
            c += (uint64_t)m * R[j] + t[j];
            t[j - 1] = (uint32_t)c;
            c >>= 32;
        }
        c += t[8];
        t[7] = (uint32_t)c;
        t[8] = (uint32_t)(c >> 32);
    }
    // t <= r here (the input is below 2^256): subtract r once if t >= r
    uint32_t d[8], borrow = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) d[j] = __builtin_subc(t[j], R[j], borrow, &borrow);
    const bool ge = t[8] || !borrow;
#pragma unroll
    for (int j = 0; j < 8; j++) w[j] = ge ? d[j] : t[j];
}

// dst[0..words) = value; a kernel rather than a memset node so the sequence
// can be captured in a HIP graph
__global__ void __launch_bounds__(256)
k_fill(uint32_t* __restrict__ dst, size_t words, uint32_t value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

// dst[0..words) = src[0..words); either side may be page-locked host memory
// mapped into the device's address space (a transfer that takes no slot in
// the copy engines' queues, so it cannot hold up copies of other streams)
__global__ void __launch_bounds__(256)
k_copy_words(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// dst[i] += src[i] (dst[i] = src[i] with first): the running entry counts of
// the point ranges of one MSM
__global__ void __launch_bounds__(256)
k_add_words(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t words, uint32_t first) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = first ? src[i] : dst[i] + src[i];
}

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
k_scatter(const job_t* __restrict__ jobs, const uint32_t* __restrict__ ebase, uint32_t njobs,
          uint32_t total_entries, const uint32_t* __restrict__ ent, const uint32_t* __restrict__ rank,
          const uint32_t* __restrict__ off, uint2* __restrict__ sorted) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total_entries) return;
    const uint32_t e = ent[t];
    if (e == kInvalid) return;
    const uint32_t j = find_owner(ebase, njobs, t);
    const uint32_t i = (t - ebase[j]) % jobs[j].n;
    const uint32_t key = e & 0x7fffffffu;
    const uint32_t pos = off[key] + rank[t];
    sorted[pos] = make_uint2(key, i | (e & 0x80000000u));
}

void launch_k_digits(dim3 grid, dim3 block, hipStream_t stream, const job_t* __restrict__ jobs, const uint32_t* __restrict__ pbase, uint32_t njobs, uint32_t total_points, uint32_t* __restrict__ ent, uint32_t* __restrict__ rank, uint32_t* __restrict__ hist) {
    hipLaunchKernelGGL(k_digits, grid, block, 0, stream, jobs, pbase, njobs, total_points, ent, rank, hist);
}

void launch_k_fill(dim3 grid, dim3 block, hipStream_t stream, uint32_t* __restrict__ dst, size_t words, uint32_t value) {
    hipLaunchKernelGGL(k_fill, grid, block, 0, stream, dst, words, value);
}

void launch_k_copy_words(hipStream_t stream, const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t words) {
    const uint32_t blocks = (uint32_t)std::min<size_t>((words + 255) / 256, 1024);
    hipLaunchKernelGGL(k_copy_words, dim3(blocks ? blocks : 1), dim3(256), 0, stream, src, dst, words);
}

void launch_k_add_words(hipStream_t stream, const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t words,
                        bool first) {
    const uint32_t blocks = (uint32_t)std::min<size_t>((words + 255) / 256, 1024);
    hipLaunchKernelGGL(k_add_words, dim3(blocks ? blocks : 1), dim3(256), 0, stream, src, dst, words, first ? 1u : 0u);
}

void launch_k_scan_tiles(dim3 grid, dim3 block, hipStream_t stream, const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t* __restrict__ tile_sums, uint32_t N) {
    hipLaunchKernelGGL(k_scan_tiles, grid, block, 0, stream, in, out, tile_sums, N);
}

void launch_k_scan_add(dim3 grid, dim3 block, hipStream_t stream, uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_offs, uint32_t N) {
    hipLaunchKernelGGL(k_scan_add, grid, block, 0, stream, out, tile_offs, N);
}

void launch_k_scatter(dim3 grid, dim3 block, hipStream_t stream, const job_t* __restrict__ jobs, const uint32_t* __restrict__ ebase, uint32_t njobs, uint32_t total_entries, const uint32_t* __restrict__ ent, const uint32_t* __restrict__ rank, const uint32_t* __restrict__ off, uint2* __restrict__ sorted) {
    hipLaunchKernelGGL(k_scatter, grid, block, 0, stream, jobs, ebase, njobs, total_entries, ent, rank, off, sorted);
}

void launch_k_count(dim3 grid, dim3 block, size_t lds, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, uint32_t* __restrict__ counts) {
    hipLaunchKernelGGL(k_count, grid, block, lds, stream, jobs, sg, tiles, counts);
}

void launch_k_prefix(dim3 grid, dim3 block, hipStream_t stream, const group_t* __restrict__ groups, const sgroup_t* __restrict__ sg, const uint32_t* __restrict__ gkb, uint32_t ngroups, uint32_t nkeys, uint32_t* __restrict__ counts, uint32_t* __restrict__ hist) {
    hipLaunchKernelGGL(k_prefix, grid, block, 0, stream, groups, sg, gkb, ngroups, nkeys, counts, hist);
}

void launch_k_place(dim3 grid, dim3 block, size_t lds, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, const uint32_t* __restrict__ counts, const uint32_t* __restrict__ off, uint2* __restrict__ sorted) {
    hipLaunchKernelGGL(k_place, grid, block, lds, stream, jobs, sg, tiles, counts, off, sorted);
}

// largest dynamic LDS the tiled sort may ask for (0 when it cannot be raised
// to what the device allows)
size_t sort_lds_limit() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_count), hipFuncAttributeMaxDynamicSharedMemorySize, v) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_place), hipFuncAttributeMaxDynamicSharedMemorySize, v) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return (size_t)v;
}

void launch_k_bin_count(dim3 grid, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, uint32_t* __restrict__ counts, uint32_t ncounts) {
    hipLaunchKernelGGL(k_bin_count, grid, dim3(kSortThreads), 0, stream, jobs, sg, tiles, counts, ncounts);
}

void launch_k_bin_place(dim3 grid, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, const uint32_t* __restrict__ pos, uint2* __restrict__ tmp) {
    hipLaunchKernelGGL(k_bin_place, grid, dim3(kSortThreads), 0, stream, jobs, sg, tiles, pos, tmp);
}

void launch_k_bin_fine(dim3 grid, hipStream_t stream, const sgroup_t* __restrict__ sg, const uint32_t* __restrict__ gkb, uint32_t ngroups, const uint32_t* __restrict__ pos, const uint2* __restrict__ tmp, uint2* __restrict__ sorted, uint32_t* __restrict__ cnt, uint32_t* __restrict__ off, uint32_t nkeys, uint32_t ncounts) {
    hipLaunchKernelGGL(k_bin_fine, grid, dim3(256), 0, stream, sg, gkb, ngroups, pos, tmp, sorted, cnt, off, nkeys, ncounts);
}

void launch_k_fr_to_scalar(hipStream_t stream, uint32_t* __restrict__ io, uint32_t n) {
    hipLaunchKernelGGL(k_fr_to_scalar, dim3((n + 255) / 256), dim3(256), 0, stream, io, n);
}

}  // namespace bls
