// Merging the partial sums of buckets that span accumulate chunks.
//
// Chunk k leaves the bucket that starts in it and continues in slot 2k+1 and
// the bucket that started before it in slot 2k (partial j = 0 of a bucket
// whose first chunk is k0 is slot 2*k0+1, partial j >= 1 is slot 2*(k0+j)).
// Most spanning buckets cross one or two boundaries: one lane adds their up
// to kLightSpan partials (a tile smaller than a bucket's entries, as for
// sizes without a tuned tile, makes three-piece buckets common; a wave each
// would cost more than the additions).  Buckets with up to kWaveSpan
// partials go to a list worked by one wave per bucket (16 quads, four-lane
// additions, strides then a shuffle tree); longer ones to a list worked by
// one workgroup per bucket (256 quads, strides then a tree in LDS).  Short top windows put many entries
// into few buckets, so a batch of MSMs can hold very many listed buckets.
#include "msm_common.hip.h"
#include "ec_quad.hip.h"
#include "launch.h"

namespace bls {

constexpr uint32_t kLightSpan = 4;
constexpr uint32_t kWaveSpan = 128;

// counts[0]: buckets listed for a wave each, counts[1]: for a workgroup each
template <class F>
__global__ void __launch_bounds__(256)
k_partial_light(const uint32_t* __restrict__ slot_key, const xyzz_t<F>* __restrict__ slots,
                const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nchunks,
                xyzz_t<F>* __restrict__ buckets, uint32_t* __restrict__ counts, uint32_t* __restrict__ wave_list,
                uint32_t* __restrict__ group_list) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nchunks) return;
    const uint32_t key = slot_key[2 * k + 1];
    if (key == kInvalid) return;
    const uint32_t m = (off[key] + cnt[key] - 1) / T - k + 1;   // partials
    if (m <= kLightSpan) {
        xyzz_t<F> a, b;
        load_xyzz(a, &slots[2 * k + 1]);
#pragma unroll 1
        for (uint32_t j = 1; j < m; j++) {
            load_xyzz(b, &slots[2 * (k + j)]);
            xyzz_add(a, a, b);
        }
        store_xyzz(&buckets[key], a);
    } else if (m <= kWaveSpan) {
        wave_list[atomicAdd(&counts[0], 1u)] = key;
    } else {
        group_list[atomicAdd(&counts[1], 1u)] = key;
    }
}

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

// One wave per listed bucket; the grid strides over the list.  Every quad
// of the wave runs every tree level (only quads below the level's width
// keep a meaningful value), so no level runs on a sparse wave.
template <class F>
__global__ void __launch_bounds__(256)
k_partial_wave(const uint32_t* __restrict__ count, const uint32_t* __restrict__ list,
               const xyzz_t<F>* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
               uint32_t T, xyzz_t<F>* __restrict__ buckets) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t q = (threadIdx.x & 63) >> 2;
    const uint32_t n = *count;
    for (uint32_t h = wave; h < n; h += nwaves) {
        const uint32_t key = list[h];
        const uint32_t k0 = off[key] / T;
        const uint32_t m = (off[key] + cnt[key] - 1) / T - k0 + 1;
        F acc, o;
        fp_zero(acc);
        for (uint32_t j = q; j < m; j += 16) {
            F x;
            quad_load(x, &slots[j == 0 ? 2 * k0 + 1 : 2 * (k0 + j)]);
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
}

// workgroup size of the long-span merge: 1024 threads for G1; the LDS tree
// of G2 points (384 bytes each) holds half as many
template <class F> struct heavy_shape { static constexpr uint32_t threads = 1024; };
template <> struct heavy_shape<fp2> { static constexpr uint32_t threads = 512; };

// One listed bucket per workgroup iteration; the grid strides over the list.
template <class F>
__global__ void __launch_bounds__(heavy_shape<F>::threads)
k_partial_heavy(const uint32_t* __restrict__ heavy_count, const uint32_t* __restrict__ heavy,
                const xyzz_t<F>* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                uint32_t T, xyzz_t<F>* __restrict__ buckets) {
    constexpr uint32_t kHeavyQuads = heavy_shape<F>::threads / 4;
    __shared__ xyzz_t<F> node[kHeavyQuads];
    const uint32_t u = threadIdx.x >> 2, ln = threadIdx.x & 3;
    const uint32_t nheavy = *heavy_count;
    for (uint32_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const uint32_t key = heavy[h];
        const uint32_t k0 = off[key] / T;
        const uint32_t m = (off[key] + cnt[key] - 1) / T - k0 + 1;   // partials
        F acc;
        fp_zero(acc);
        for (uint32_t j = u; j < m; j += kHeavyQuads) {
            F x;
            quad_load(x, &slots[j == 0 ? 2 * k0 + 1 : 2 * (k0 + j)]);
            quad_add(acc, acc, x);
        }
        reinterpret_cast<F*>(&node[u])[ln] = acc;
        __syncthreads();
        // pairwise tree over the quads that hold partials; at least four
        // quads (16 lanes) run each level, the extra ones repeating a merge
        // without storing it
        uint32_t width = 1;
        while (width < min(m, kHeavyQuads)) width <<= 1;
        for (uint32_t half = width / 2; half >= 1; half >>= 1) {
            F r;
            const bool run = u < max(half, 4u);
            if (run) {
                const uint32_t v = u % half;
                F a, b;
                a = reinterpret_cast<const F*>(&node[v])[ln];
                b = reinterpret_cast<const F*>(&node[v + half])[ln];
                quad_add(r, a, b);
            }
            __syncthreads();
            if (u < half) reinterpret_cast<F*>(&node[u])[ln] = r;
            __syncthreads();
        }
        if (u == 0) quad_store(&buckets[key], reinterpret_cast<const F*>(&node[0])[ln]);
        __syncthreads();
    }
}

template <class F>
void launch_k_partial_light(dim3 grid, dim3 block, hipStream_t stream, const uint32_t* __restrict__ slot_key, const xyzz_t<F>* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nchunks, xyzz_t<F>* __restrict__ buckets, uint32_t* __restrict__ counts, uint32_t* __restrict__ wave_list, uint32_t* __restrict__ group_list) {
    hipLaunchKernelGGL(k_partial_light<F>, grid, block, 0, stream, slot_key, slots, off, cnt, T, nchunks, buckets, counts, wave_list, group_list);
}

template <class F>
void launch_k_partial_wave(dim3 grid, hipStream_t stream, const uint32_t* __restrict__ count, const uint32_t* __restrict__ list, const xyzz_t<F>* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, xyzz_t<F>* __restrict__ buckets) {
    hipLaunchKernelGGL(k_partial_wave<F>, grid, dim3(256), 0, stream, count, list, slots, off, cnt, T, buckets);
}

template <class F>
void launch_k_partial_heavy(dim3 grid, hipStream_t stream, const uint32_t* __restrict__ heavy_count, const uint32_t* __restrict__ heavy, const xyzz_t<F>* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, xyzz_t<F>* __restrict__ buckets) {
    hipLaunchKernelGGL(k_partial_heavy<F>, grid, dim3(heavy_shape<F>::threads), 0, stream, heavy_count, heavy, slots, off, cnt, T, buckets);
}

#define BLS_PARTIAL_INSTANCES(F)                                                                                   \
    template void launch_k_partial_light<F>(dim3, dim3, hipStream_t, const uint32_t*, const xyzz_t<F>*,          \
                                            const uint32_t*, const uint32_t*, uint32_t, uint32_t, xyzz_t<F>*,    \
                                            uint32_t*, uint32_t*, uint32_t*);                                    \
    template void launch_k_partial_wave<F>(dim3, hipStream_t, const uint32_t*, const uint32_t*, const xyzz_t<F>*, \
                                           const uint32_t*, const uint32_t*, uint32_t, xyzz_t<F>*);               \
    template void launch_k_partial_heavy<F>(dim3, hipStream_t, const uint32_t*, const uint32_t*,                  \
                                            const xyzz_t<F>*, const uint32_t*, const uint32_t*, uint32_t,         \
                                            xyzz_t<F>*);
BLS_PARTIAL_INSTANCES(fp)
BLS_PARTIAL_INSTANCES(fp2)

}  // namespace bls
