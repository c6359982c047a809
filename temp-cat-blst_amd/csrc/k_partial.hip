// Merging the partial sums of buckets that span accumulate chunks.
//
// Chunk k leaves the bucket that starts in it and continues in slot 2k+1 and
// the bucket that started before it in slot 2k (partial j = 0 of a bucket
// whose first chunk is k0 is slot 2*k0+1, partial j >= 1 is slot 2*(k0+j)).
// Most spanning buckets cross one boundary: two pieces, one addition, one
// lane per chunk.  Buckets of more pieces were listed by the accumulate
// kernel by piece count (a tile smaller than a bucket's entries, as for sizes
// without a tuned tile, makes three- and four-piece buckets common; short top
// windows and repeated scalars put very many entries into few buckets).  One
// launch runs every kind at once, in block ranges of its own, so the
// latency-bound merges of the longest buckets overlap the others instead of
// following them:
//   heavy  (first blocks) one workgroup per bucket of more than kWaveSpan
//          pieces: 64 quads stride over the pieces, then a tree in LDS;
//   wave   one wave per bucket of up to kWaveSpan pieces (16 quads, strides
//          then a shuffle tree);
//   few    one lane per bucket of three or four pieces;
//   pairs  one lane per chunk whose bucket has exactly two pieces.
#include "msm_common.hip.h"
#include "ec_quad.hip.h"
#include "launch.h"

namespace bls {

constexpr uint32_t kMergeThreads = 256;
constexpr uint32_t kHeavyQuads = kMergeThreads / 4;
// the fewest pieces a part of a split bucket holds: eight per quad
constexpr uint32_t kPartPieces = 8 * kHeavyQuads;

template <int D>
__device__ __forceinline__ void fp_shfl_down(fp& r, const fp& a) {
#pragma unroll
    for (int k = 0; k < 12; k++) r.l[k] = (uint32_t)__shfl_down((int)a.l[k], D, 64);
}

template <int D>
__device__ __forceinline__ void fp_shfl_down(fp2& r, const fp2& a) {
    fp_shfl_down<D>(r.c0, a.c0);
    fp_shfl_down<D>(r.c1, a.c1);
}

// pieces 0..m-1 of the bucket whose first chunk is k0
__device__ __forceinline__ uint32_t piece_slot(uint32_t k0, uint32_t j) { return j == 0 ? 2 * k0 + 1 : 2 * (k0 + j); }

// Pairwise tree over the block's quads in LDS: r is the sum of every quad's
// acc, on every quad.  Every thread of the block must call it.
template <class F>
__device__ __forceinline__ void heavy_tree(F& r, xyzz_t<F>* node, F acc, uint32_t u, uint32_t ln) {
    reinterpret_cast<F*>(&node[u])[ln] = acc;
    __syncthreads();
    // at least four quads (16 lanes) run each level, the extra ones
    // repeating a merge unstored
    for (uint32_t half = kHeavyQuads / 2; half >= 1; half >>= 1) {
        F x;
        if (u < max(half, 4u)) {
            const uint32_t v = u % half;
            F y = reinterpret_cast<const F*>(&node[v])[ln];
            F z = reinterpret_cast<const F*>(&node[v + half])[ln];
            quad_add(x, y, z);
        }
        __syncthreads();
        if (u < half) reinterpret_cast<F*>(&node[u])[ln] = x;
        __syncthreads();
    }
    r = reinterpret_cast<const F*>(&node[0])[ln];
    __syncthreads();
}

// sum over the block of the pieces j0 <= j < j1 of the bucket whose first
// chunk is k0 (quads stride over them, then the tree)
template <class F>
__device__ __forceinline__ void heavy_sum(F& r, xyzz_t<F>* node, const xyzz_t<F>* slots, uint32_t k0, uint32_t j0,
                                          uint32_t j1, uint32_t u, uint32_t ln) {
    F acc;
    fp_zero(acc);
    for (uint32_t j = j0 + u; j < j1; j += kHeavyQuads) {
        F x;
        quad_load(x, &slots[piece_slot(k0, j)]);
        quad_add(acc, acc, x);
    }
    heavy_tree(r, node, acc, u, ln);
}

// the G part sums of a bucket of m pieces, part s' stored at piece m s' / G
template <class F>
__device__ __forceinline__ void heavy_sum_parts(F& r, xyzz_t<F>* node, const xyzz_t<F>* slots, uint32_t k0,
                                                uint32_t m, uint32_t G, uint32_t u, uint32_t ln) {
    F acc;
    fp_zero(acc);
    for (uint32_t s = u; s < G; s += kHeavyQuads) {
        F x;
        quad_load(x, &slots[piece_slot(k0, (uint32_t)((uint64_t)m * s / G))]);
        quad_add(acc, acc, x);
    }
    heavy_tree(r, node, acc, u, ln);
}

// block ranges of the roles: heavy, wave, few, pairs
struct merge_blocks_t {
    uint32_t heavy, wave, few, pairs;
};

template <class F>
__global__ void __launch_bounds__(kMergeThreads)
k_partial_merge(const uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots,
                const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nchunks,
                xyzz_t<F>* __restrict__ buckets, merge_lists_t lists, merge_blocks_t nb) {
    __shared__ xyzz_t<F> node[kHeavyQuads];
    __shared__ uint32_t last_part;
    uint32_t b = blockIdx.x;
    if (b < nb.heavy) {
        // Listed buckets (more than kWaveSpan pieces) over the role's blocks.
        // With fewer buckets than blocks, each bucket has G blocks, and
        // bucket h's m pieces are cut into Gh = min(G, m / kPartPieces)
        // parts (so each quad still adds eight pieces or more in a row; a
        // part's second tree and fence cost more than shorter rows save).
        // Part s sums pieces [m s / Gh, m (s + 1) / Gh) into the slot of its
        // first piece, and the block that finishes a bucket's last part adds
        // the Gh part sums.
        const uint32_t u = threadIdx.x >> 2, ln = threadIdx.x & 3;
        const uint32_t n = lists.counts[1];
        const uint32_t G = (n != 0 && n < nb.heavy) ? min(nb.heavy / n, kHeavyParts) : 1u;
        for (uint32_t item = b; item < n * G; item += nb.heavy) {
            const uint32_t h = item / G, s = item - h * G;
            const uint32_t key = lists.group[h];
            const uint32_t k0 = off[key] / T;
            const uint32_t m = (off[key] + cnt[key] - 1) / T - k0 + 1;
            const uint32_t Gh = min(G, max(m / kPartPieces, 1u));
            if (s >= Gh) continue;   // uniform over the block
            const uint32_t j0 = (uint32_t)((uint64_t)m * s / Gh), j1 = (uint32_t)((uint64_t)m * (s + 1) / Gh);
            F r;
            heavy_sum(r, node, slots, k0, j0, j1, u, ln);
            if (Gh == 1) {
                if (u == 0) quad_store(&buckets[key], r);
            } else {
                if (u == 0) quad_store(&slots[piece_slot(k0, j0)], r);
                __threadfence();   // the part sum is visible device-wide before it is counted
                __syncthreads();
                if (threadIdx.x == 0) {
                    last_part = atomicAdd(&lists.parts[h], 1u) == Gh - 1;
                    __threadfence();
                }
                __syncthreads();
                if (last_part) {
                    __threadfence();   // every other part's sum is read after its count
                    // part s' starts at piece m s' / Gh: parts as pieces at those indices
                    heavy_sum_parts(r, node, slots, k0, m, Gh, u, ln);
                    if (u == 0) quad_store(&buckets[key], r);
                }
            }
            __syncthreads();
        }
        return;
    }
    b -= nb.heavy;
    if (b < nb.wave) {
        // every quad of a wave runs every shuffle level (only quads below the
        // level's width keep a meaningful value), so no level runs sparse
        const uint32_t wave = (b * kMergeThreads + threadIdx.x) >> 6;
        const uint32_t nwaves = (nb.wave * kMergeThreads) >> 6;
        const uint32_t q = (threadIdx.x & 63) >> 2;
        const uint32_t n = lists.counts[0];
        for (uint32_t h = wave; h < n; h += nwaves) {
            const uint32_t key = lists.wave[h];
            const uint32_t k0 = off[key] / T;
            const uint32_t m = (off[key] + cnt[key] - 1) / T - k0 + 1;
            F acc, o;
            fp_zero(acc);
            for (uint32_t j = q; j < m; j += 16) {
                F x;
                quad_load(x, &slots[piece_slot(k0, j)]);
                quad_add(acc, acc, x);
            }
            fp_shfl_down<32>(o, acc);
            quad_add(acc, acc, o);
            fp_shfl_down<16>(o, acc);
            quad_add(acc, acc, o);
            fp_shfl_down<8>(o, acc);
            quad_add(acc, acc, o);
            fp_shfl_down<4>(o, acc);
            quad_add(acc, acc, o);
            if (q == 0) quad_store(&buckets[key], acc);
        }
        return;
    }
    b -= nb.wave;
    if (b < nb.few) {
        const uint32_t n = lists.counts[2];
        for (uint32_t h = b * kMergeThreads + threadIdx.x; h < n; h += nb.few * kMergeThreads) {
            const uint32_t key = lists.few[h];
            const uint32_t k0 = off[key] / T;
            const uint32_t m = (off[key] + cnt[key] - 1) / T - k0 + 1;
            xyzz_t<F> a, c;
            load_xyzz(a, &slots[2 * k0 + 1]);
#pragma unroll 1
            for (uint32_t j = 1; j < m; j++) {
                load_xyzz(c, &slots[2 * (k0 + j)]);
                xyzz_add(a, a, c);
            }
            store_xyzz(&buckets[key], a);
        }
        return;
    }
    b -= nb.few;
    const uint32_t k = b * kMergeThreads + threadIdx.x;
    if (k >= nchunks) return;
    const uint32_t key = slot_key[2 * k + 1];
    if (key == kInvalid) return;
    if ((off[key] + cnt[key] - 1) / T != k + 1) return;   // not exactly two pieces: listed
    xyzz_t<F> a, c;
    load_xyzz(a, &slots[2 * k + 1]);
    load_xyzz(c, &slots[2 * k + 2]);
    xyzz_add(a, a, c);
    store_xyzz(&buckets[key], a);
}

template <class F>
void launch_k_partial_merge(hipStream_t stream, const uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nchunks, xyzz_t<F>* __restrict__ buckets, merge_lists_t lists) {
    const uint32_t per = (nchunks + kMergeThreads - 1) / kMergeThreads;
    merge_blocks_t nb;
    nb.heavy = std::min<uint32_t>(per, 256);
    nb.wave = std::min<uint32_t>((nchunks + 63) / 64, 512);
    nb.few = per;
    nb.pairs = per;
    hipLaunchKernelGGL(k_partial_merge<F>, dim3(nb.heavy + nb.wave + nb.few + nb.pairs), dim3(kMergeThreads), 0, stream,
                       slot_key, slots, off, cnt, T, nchunks, buckets, lists, nb);
}

#define BLS_PARTIAL_INSTANCES(F)                                                                                   \
    template void launch_k_partial_merge<F>(hipStream_t, const uint32_t*, xyzz_t<F>*, const uint32_t*,     \
                                            const uint32_t*, uint32_t, uint32_t, xyzz_t<F>*, merge_lists_t);
BLS_PARTIAL_INSTANCES(fp)
BLS_PARTIAL_INSTANCES(fp2)

}  // namespace bls
