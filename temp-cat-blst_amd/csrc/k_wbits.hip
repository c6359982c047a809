// Fixed-base tables (blst_p1s_mult_wbits_precompute / blst_p1s_mult_wbits and
// their G2 forms, blst/scalar/curve/multi_scalar.cpp:268-286, 481-707).  The
// row, sum and tree kernels are templates over the field (fp: G1, fp2: G2);
// the comb kernels below them are G1 only.
//
// A table holds, for each point P_i, the row [1]P_i ... [m]P_i (m =
// 2^(wbits-1)) in affine form.  The MSM over a table needs no buckets: with
// the Booth digits d_{i,w} of the scalars (windows of wbits bits),
//   sum_i k_i P_i = sum_w 2^(wbits w) S_w,   S_w = sum_i sign(d) row_i[|d| - 1],
// so every (window, point) pair is one gathered entry and one mixed
// addition into its window's sum.
//
//   rows        L lanes per row, each lane one run of m / L consecutive
//               multiples: [t (m/L) + 1]P by double-and-add, then + P per
//               entry (XYZZ); k_to_affine then normalises the rows
//   sum         one lane per (window, run of C points): gathers and mixed
//               additions into one XYZZ partial sum
//   tree        one workgroup per window: 256 quads fold the window's
//               partials (four-lane additions, strides then a tree in LDS);
//               the window sum is written canonical for the host's fold
#include "msm_common.hip.h"
#include "ec_quad.hip.h"
#include "launch.h"

namespace bls {

__device__ __forceinline__ void load_aff_at(aff2& a, const uint4* __restrict__ base, size_t idx) {
    const uint4* s = base + idx * 12;
    fp_from_uint4(a.x.c0, s);
    fp_from_uint4(a.x.c1, s + 3);
    fp_from_uint4(a.y.c0, s + 6);
    fp_from_uint4(a.y.c1, s + 9);
}

__device__ __forceinline__ void load_aff_at(aff& a, const uint4* __restrict__ base, size_t idx) {
    const uint4* s = base + idx * 6;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint4 u = s[k], v = s[k + 3];
        a.x.l[4 * k] = u.x; a.x.l[4 * k + 1] = u.y; a.x.l[4 * k + 2] = u.z; a.x.l[4 * k + 3] = u.w;
        a.y.l[4 * k] = v.x; a.y.l[4 * k + 1] = v.y; a.y.l[4 * k + 2] = v.z; a.y.l[4 * k + 3] = v.w;
    }
}

// rows [row0, row0 + nrows) of the table, as XYZZ multiples: out[r * m + j]
// = [j + 1] P_{row0 + r}; all-zero points give infinity entries.
template <class F>
__global__ void __launch_bounds__(256)
k_wbits_rows(const uint4* __restrict__ pts, uint32_t row0, uint32_t nrows, uint32_t m, uint32_t L,
             xyzz_t<F>* __restrict__ out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrows * L) return;
    const uint32_t r = g / L, t = g % L, S = m / L;
    aff_t<F> P;
    load_aff_at(P, pts, (size_t)row0 + r);
    xyzz_t<F>* dst = out + (size_t)r * m + (size_t)t * S;
    xyzz_t<F> acc;
    xyzz_set_inf(acc);
    if (aff_is_inf(P)) {
        for (uint32_t s = 0; s < S; s++) store_xyzz(&dst[s], acc);
        return;
    }
    // acc = [t S + 1] P, most significant bit first
    const uint32_t k = t * S + 1;
    for (int b = 31 - __builtin_clz(k); b >= 0; b--) {
        xyzz_dbl_inl(acc, acc);
        if ((k >> b) & 1u) xyzz_madd(acc, P);
    }
    store_xyzz(&dst[0], acc);
    for (uint32_t s = 1; s < S; s++) {
        xyzz_madd(acc, P);
        store_xyzz(&dst[s], acc);
    }
}

// Partial window sums.  Lane (w, c) adds the entries of points
// [c C, min((c + 1) C, n)) in window w into part[w * nch + c].  With a table
// (device memory, n rows of m entries) the lane computes the Booth digit of
// each scalar itself; without one, list holds the D * n entries already
// gathered and signed (window-major, all-zero = nothing to add).
template <class F>
__global__ void __launch_bounds__(256)
k_wbits_sum(const uint4* __restrict__ table, const uint8_t* __restrict__ scalars, uint32_t nbytes,
            uint32_t nbits, uint32_t vec, const uint4* __restrict__ list, uint32_t n, uint32_t wbits, uint32_t D,
            uint32_t C, uint32_t nch, xyzz_t<F>* __restrict__ part) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= D * nch) return;
    const uint32_t w = g / nch, c = g % nch;
    const uint32_t i0 = c * C, i1 = min(i0 + C, n);
    const size_t m = (size_t)1 << (wbits - 1);
    xyzz_t<F> acc;
    xyzz_set_inf(acc);
    for (uint32_t i = i0; i < i1; i++) {
        aff_t<F> a;
        if (table) {
            const int32_t d = booth_digit_at(scalars + (size_t)i * nbytes, vec != 0, w * wbits, wbits, nbits);
            if (d == 0) continue;
            load_aff_at(a, table, (size_t)i * m + (size_t)(d < 0 ? -d : d) - 1);
            if (aff_is_inf(a)) continue;
            if (d < 0) fp_neg(a.y, a.y);
        } else {
            load_aff_at(a, list, (size_t)w * n + i);
            if (aff_is_inf(a)) continue;
        }
        xyzz_madd(acc, a);
    }
    store_xyzz(&part[g], acc);
}

// Window w's nch partials -> W[w] (canonical XYZZ); one workgroup of
// THREADS per window (1024 for G1, 512 for G2, whose LDS nodes are twice as
// large).
template <class F, uint32_t THREADS>
__global__ void __launch_bounds__(THREADS)
k_wbits_tree(const xyzz_t<F>* __restrict__ part, uint32_t nch, xyzz_t<F>* __restrict__ W) {
    constexpr uint32_t kQuads = THREADS / 4;
    __shared__ xyzz_t<F> node[kQuads];
    const uint32_t w = blockIdx.x;
    const uint32_t u = threadIdx.x >> 2, ln = threadIdx.x & 3;
    const xyzz_t<F>* src = part + (size_t)w * nch;
    F acc;
    fp_zero(acc);
    for (uint32_t j = u; j < nch; j += kQuads) {
        F x;
        quad_load(x, &src[j]);
        quad_add(acc, acc, x);
    }
    reinterpret_cast<F*>(&node[u])[ln] = acc;
    __syncthreads();
    // pairwise tree over the quads that hold partials; at least four quads
    // (16 lanes) run each level, the extra ones repeating a merge without
    // storing it
    uint32_t width = 1;
    while (width < min(nch, kQuads)) width <<= 1;
    for (uint32_t half = width / 2; half >= 1; half >>= 1) {
        F r;
        if (u < max(half, 4u)) {
            const uint32_t v = u % half;
            const F a = reinterpret_cast<const F*>(&node[v])[ln];
            const F b = reinterpret_cast<const F*>(&node[v + half])[ln];
            quad_add(r, a, b);
        }
        __syncthreads();
        if (u < half) reinterpret_cast<F*>(&node[u])[ln] = r;
        __syncthreads();
    }
    if (u == 0) {
        F r = reinterpret_cast<const F*>(&node[0])[ln];
        fp_canon(r, r);
        quad_store(&W[w], r);
    }
}

// Comb tables of a resident setup: window w of the table holds the rows of
// Q_i = [2^(wbits w)] P_i, so sum_i k_i P_i = sum_(w, i) sign(d) row_(w, i)[|d| - 1]
// with no doublings between windows.  k_shift_points computes the next
// window's bases: out_i = [2^s] in_i (XYZZ, then k_to_affine).
__global__ void __launch_bounds__(256)
k_shift_points(const uint4* __restrict__ in, uint32_t n, uint32_t s, xyzz* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aff a;
    load_aff_at(a, in, i);
    xyzz acc;
    xyzz_set_inf(acc);
    if (!aff_is_inf(a)) {
        xyzz_from_aff(acc, a);
#pragma unroll 1
        for (uint32_t k = 0; k < s; k++) xyzz_dbl_inl(acc, acc);
    }
    store_xyzz(&out[i], acc);
}

// Batched sums over a comb table: lane (job, c) adds the terms t in
// [c C, min((c + 1) C, D n)) of its job, term t = w n + i being the entry of
// point i's Booth digit in window w, into part[job * nch + c].  Scalars are
// job-major (job j's n scalars at scalars + j n nbytes).
__global__ void __launch_bounds__(256)
k_comb_sum(const uint4* __restrict__ table, const uint8_t* __restrict__ scalars, uint32_t nbytes, uint32_t nbits,
           uint32_t vec, uint32_t n, uint32_t wbits, uint32_t D, uint32_t njobs, uint32_t C, uint32_t nch,
           xyzz* __restrict__ part) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= njobs * nch) return;
    const uint32_t job = g / nch, c = g % nch;
    const uint32_t terms = D * n;
    const uint32_t t0 = c * C, t1 = min(t0 + C, terms);
    const size_t m = (size_t)1 << (wbits - 1);
    const uint8_t* sc = scalars + (size_t)job * n * nbytes;
    xyzz acc;
    xyzz_set_inf(acc);
    uint32_t w = t0 / n, i = t0 % n;
    for (uint32_t t = t0; t < t1; t++) {
        const int32_t d = booth_digit_at(sc + (size_t)i * nbytes, vec != 0, w * wbits, wbits, nbits);
        if (d != 0) {
            aff a;
            load_aff_at(a, table, ((size_t)w * n + i) * m + (size_t)(d < 0 ? -d : d) - 1);
            if (!aff_is_inf(a)) {
                if (d < 0) fp_neg(a.y, a.y);
                xyzz_madd(acc, a);
            }
        }
        if (++i == n) { i = 0; w++; }
    }
    store_xyzz(&part[g], acc);
}

void launch_k_shift_points(hipStream_t stream, const uint4* __restrict__ in, uint32_t n, uint32_t s,
                           xyzz* __restrict__ out) {
    hipLaunchKernelGGL(k_shift_points, dim3((n + 255) / 256), dim3(256), 0, stream, in, n, s, out);
}

void launch_k_comb_sum(hipStream_t stream, const uint4* __restrict__ table, const uint8_t* __restrict__ scalars,
                       uint32_t nbytes, uint32_t nbits, uint32_t vec, uint32_t n, uint32_t wbits, uint32_t D,
                       uint32_t njobs, uint32_t C, uint32_t nch, xyzz* __restrict__ part) {
    const uint32_t lanes = njobs * nch;
    hipLaunchKernelGGL(k_comb_sum, dim3((lanes + 255) / 256), dim3(256), 0, stream, table, scalars, nbytes, nbits,
                       vec, n, wbits, D, njobs, C, nch, part);
}

template <class F>
void launch_k_wbits_rows(hipStream_t stream, const uint4* __restrict__ pts, uint32_t row0, uint32_t nrows, uint32_t m,
                         uint32_t L, xyzz_t<F>* __restrict__ out) {
    const uint32_t lanes = nrows * L;
    hipLaunchKernelGGL(k_wbits_rows<F>, dim3((lanes + 255) / 256), dim3(256), 0, stream, pts, row0, nrows, m, L, out);
}

template <class F>
void launch_k_wbits_sum(hipStream_t stream, const uint4* __restrict__ table, const uint8_t* __restrict__ scalars,
                        uint32_t nbytes, uint32_t nbits, uint32_t vec, const uint4* __restrict__ list, uint32_t n,
                        uint32_t wbits, uint32_t D, uint32_t C, uint32_t nch, xyzz_t<F>* __restrict__ part) {
    const uint32_t lanes = D * nch;
    hipLaunchKernelGGL(k_wbits_sum<F>, dim3((lanes + 255) / 256), dim3(256), 0, stream, table, scalars, nbytes, nbits,
                       vec, list, n, wbits, D, C, nch, part);
}

template <class F>
void launch_k_wbits_tree(hipStream_t stream, const xyzz_t<F>* __restrict__ part, uint32_t nch, uint32_t D,
                         xyzz_t<F>* __restrict__ W) {
    constexpr uint32_t threads = sizeof(F) == sizeof(fp) ? 1024 : 512;
    hipLaunchKernelGGL((k_wbits_tree<F, threads>), dim3(D), dim3(threads), 0, stream, part, nch, W);
}

#define BLS_WBITS_INSTANCES(F)                                                                                      \
    template void launch_k_wbits_rows<F>(hipStream_t, const uint4*, uint32_t, uint32_t, uint32_t, uint32_t,        \
                                         xyzz_t<F>*);                                                               \
    template void launch_k_wbits_sum<F>(hipStream_t, const uint4*, const uint8_t*, uint32_t, uint32_t, uint32_t,   \
                                        const uint4*, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, xyzz_t<F>*); \
    template void launch_k_wbits_tree<F>(hipStream_t, const xyzz_t<F>*, uint32_t, uint32_t, xyzz_t<F>*);
BLS_WBITS_INSTANCES(fp)
BLS_WBITS_INSTANCES(fp2)

}  // namespace bls
