// Measured probe of batch-affine addition against the XYZZ mixed addition
// of the bucket accumulation, on the same point pairs (test kit only).
//
// Batch-affine (the CPU algorithm of blst/scalar/curve/multi_scalar_batch_affine.cpp:267-345
// and bulk_addition.h:116-210): P + Q = (x3, y3) with
// lambda = (y2 - y1)/(x2 - x1), x3 = lambda^2 - x1 - x2, y3 = lambda(x1 - x3) - y1,
// every denominator of a batch inverted at once by Montgomery's trick.  Here
// a batch is one 256-lane workgroup of K pairs per lane, the most favourable
// arrangement this GPU allows (no bucket bookkeeping, no exceptional inputs):
//   forward   per lane: d_k = x2 - x1, prefix products pi_k = d_0 ... d_k
//             (stored to memory: K x 48 bytes per lane);
//   workgroup the 256 lane totals through LDS to wave 0: four per lane, an
//             in-wave prefix and suffix scan by shuffles, ONE divstep
//             inversion of the workgroup product, and the inverse of every
//             lane total back through LDS;
//   backward  per lane, k = K-1 .. 0: 1/d_k = inv * pi_(k-1), inv *= d_k,
//             then lambda, x3, y3.
// Six products per addition (1 forward, 2 backward, lambda, lambda^2, y3)
// against ten in madd-2008-s, plus the workgroup's share.
//
// The XYZZ probe runs acc = P (ZZ = ZZZ = 1), acc += Q with xyzz_madd: the
// per-entry work of k_accumulate.
//
// Pairs: pair j of the grid is (P_a, P_b) with a = 2j, b = 2j + 1, or the
// points idx[2j], idx[2j + 1] when an index is given (the random gathers of
// a bucket list).  Lanes take pairs j = k * G + g (G lanes in all) so that
// the lanes of a wave read consecutive pairs.
#include "msm_common.hip.h"
#include "fq.hip.h"

namespace bls {

constexpr uint32_t kProbeThreads = 256;

__device__ __forceinline__ void fp_shfl(fp& r, const fp& a, int src) {
#pragma unroll
    for (int k = 0; k < 12; k++) r.l[k] = (uint32_t)__shfl((int)a.l[k], src, 64);
}

__device__ __forceinline__ void fp_set(fp& r, bool c, const fp& a) {
#pragma unroll
    for (int k = 0; k < 12; k++) r.l[k] = c ? a.l[k] : r.l[k];
}

__device__ __forceinline__ void load_fp(fp& r, const fp* __restrict__ src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const uint4 v = s[q];
        r.l[4 * q] = v.x; r.l[4 * q + 1] = v.y; r.l[4 * q + 2] = v.z; r.l[4 * q + 3] = v.w;
    }
}

__device__ __forceinline__ void store_fp(fp* __restrict__ dst, const fp& a) {
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int q = 0; q < 3; q++) d[q] = make_uint4(a.l[4 * q], a.l[4 * q + 1], a.l[4 * q + 2], a.l[4 * q + 3]);
}

__device__ __forceinline__ void pair_points(uint32_t j, const uint32_t* __restrict__ idx, uint32_t& a, uint32_t& b) {
    if (idx) {
        a = idx[2 * j];
        b = idx[2 * j + 1];
    } else {
        a = 2 * j;
        b = 2 * j + 1;
    }
}

// 1 where x2 == x1 (the pair is exceptional: doubling or P + (-P)); the
// probe then uses d = 1 and leaves the pair's output to the caller's check
__device__ __forceinline__ void pair_denominator(fp& d, const fp& x1, const fp& x2) {
    fp_sub(d, x2, x1);
    if (fp_is_zero(d)) fp_one(d);
}

// three waves per SIMD, as the accumulation runs: the one inversion per
// workgroup (wave 0 only) may spill, the per-pair loops do not
template <int K>
__global__ void __launch_bounds__(kProbeThreads, 3)
k_ba_pairs(const uint4* __restrict__ pts, const uint32_t* __restrict__ idx, fp* __restrict__ pi,
           aff* __restrict__ out, uint32_t G) {
    __shared__ fp tot[kProbeThreads];
    // wave 0's stash: partial products and scans, so that few field elements
    // are live at once (the inversion's own state is large)
    __shared__ fp st_p1[64], st_p2[64], st_s[64], st_u[64];
    const uint32_t t = threadIdx.x;
    const uint32_t g = blockIdx.x * kProbeThreads + t;
    // forward: prefix products of the denominators (x coordinates only)
    fp acc;
#pragma unroll 1
    for (int k = 0; k < K; k++) {
        uint32_t ia, ib;
        pair_points((uint32_t)k * G + g, idx, ia, ib);
        fp x1, x2;
        load_fp(x1, reinterpret_cast<const fp*>(pts + (size_t)ia * 6));
        load_fp(x2, reinterpret_cast<const fp*>(pts + (size_t)ib * 6));
        fp d;
        pair_denominator(d, x1, x2);
        if (k == 0) acc = d;
        else fp_mul(acc, acc, d);
        store_fp(&pi[(size_t)k * G + g], acc);
    }
    tot[t] = acc;
    __syncthreads();
    if (t < 64) {
        // lane j of wave 0 owns totals 4j .. 4j+3: p1 = a0 a1, p2 = p1 a2,
        // p3 = p2 a3
        fp p, q;
        fp_mul(p, tot[4 * t], tot[4 * t + 1]);
        st_p1[t] = p;
        fp_mul(p, p, tot[4 * t + 2]);
        st_p2[t] = p;
        fp_mul(p, p, tot[4 * t + 3]);
        // inclusive prefix s_j and suffix u_j products of p3 over the wave
        q = p;
#pragma unroll 1
        for (int off = 1; off < 64; off <<= 1) {
            fp x, xs;
            fp_shfl(x, p, (int)t - off);
            fp_mul(xs, x, p);
            fp_set(p, (int)t >= off, xs);
        }
        st_s[t] = p;
#pragma unroll 1
        for (int off = 1; off < 64; off <<= 1) {
            fp y, yu;
            fp_shfl(y, q, (int)t + off);
            fp_mul(yu, q, y);
            fp_set(q, (int)t + off < 64, yu);
        }
        st_u[t] = q;
        // one inversion of the workgroup's product s_63
        fp all, w;
        fp_shfl(all, p, 63);
        fp_inv(w, all);
        // 1/p3_j = w * s_(j-1) * u_(j+1)
        fp e, f, one;
        fp_one(one);
        p = st_s[t];
        q = st_u[t];
        fp_shfl(e, p, (int)t - 1);
        fp_shfl(f, q, (int)t + 1);
        fp_set(e, t == 0, one);
        fp_set(f, t == 63, one);
        fp_mul(w, w, e);
        fp_mul(w, w, f);
        // the four inverses of the lane's totals, from 1/p3
        fp r;
        fp_mul(r, w, st_p2[t]);
        const fp a3 = tot[4 * t + 3];
        fp_mul(w, w, a3);                    // 1/p2
        tot[4 * t + 3] = r;
        fp_mul(r, w, st_p1[t]);
        const fp a2 = tot[4 * t + 2];
        fp_mul(w, w, a2);                    // 1/p1
        tot[4 * t + 2] = r;
        const fp a0 = tot[4 * t], a1 = tot[4 * t + 1];
        fp_mul(r, w, a0);
        tot[4 * t + 1] = r;
        fp_mul(r, w, a1);
        tot[4 * t] = r;
    }
    __syncthreads();
    fp inv = tot[t];   // 1 / pi_(K-1)
#pragma unroll 1
    for (int k = K - 1; k >= 0; k--) {
        const uint32_t j = (uint32_t)k * G + g;
        uint32_t ia, ib;
        pair_points(j, idx, ia, ib);
        aff P, Q;
        load_aff(P, pts, ia);
        load_aff(Q, pts, ib);
        fp dinv;
        if (k > 0) {
            fp prev, d;
            load_fp(prev, &pi[(size_t)(k - 1) * G + g]);
            fp_mul(dinv, inv, prev);
            pair_denominator(d, P.x, Q.x);
            fp_mul(inv, inv, d);
        } else {
            dinv = inv;
        }
        fp lam, t1, x3, y3;
        fp_sub(t1, Q.y, P.y);
        fp_mul(lam, t1, dinv);
        fp_sqr(x3, lam);
        fp_sub(x3, x3, P.x);
        fp_sub(x3, x3, Q.x);
        fp_sub(t1, P.x, x3);
        fp_mul(y3, lam, t1);
        fp_sub(y3, y3, P.y);
        aff o;
        fp_canon(o.x, x3);
        fp_canon(o.y, y3);
        store_fp(&out[j].x, o.x);
        store_fp(&out[j].y, o.y);
    }
}

template <int K>
__global__ void __launch_bounds__(kProbeThreads)
k_madd_pairs(const uint4* __restrict__ pts, const uint32_t* __restrict__ idx, xyzz* __restrict__ out, uint32_t G) {
    const uint32_t g = blockIdx.x * kProbeThreads + threadIdx.x;
#pragma unroll 1
    for (int k = 0; k < K; k++) {
        const uint32_t j = (uint32_t)k * G + g;
        uint32_t ia, ib;
        pair_points(j, idx, ia, ib);
        aff P, Q;
        load_aff(P, pts, ia);
        load_aff(Q, pts, ib);
        xyzz acc;
        xyzz_from_aff(acc, P);
        xyzz_madd(acc, Q);
        xyzz c;
        xyzz_canon(c, acc);
        store_xyzz(&out[j], c);
    }
}

// The memory side of the bucket accumulation: lane g folds the points
// idx[k G + g], k < kGatherK, into one bucket sum with the accumulation
// kernel's arithmetic (fq_xyzz_madd), reading them at a stride of S16 uint4
// (6: the 96-byte blst_p1_affine array; 8: the same points padded to 128
// bytes, one cache line each), and stores the sum's word image.
constexpr uint32_t kGatherK = 32;

template <int S16>
__global__ void __launch_bounds__(kProbeThreads, 3)
k_gather_sums(const uint4* __restrict__ pts, const uint32_t* __restrict__ idx, xyzz* __restrict__ out, uint32_t G) {
    const uint32_t g = blockIdx.x * kProbeThreads + threadIdx.x;
    fq_xyzz acc;
    fq_xyzz_set_inf(acc);
#pragma unroll 1
    for (uint32_t k = 0; k < kGatherK; k++) {
        const uint4* s = pts + (size_t)idx[k * G + g] * S16;
        uint32_t w[24];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const uint4 v = s[q];
            w[4 * q] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
        }
        uint32_t z = 0;
#pragma unroll
        for (int i = 0; i < 24; i++) z |= w[i];
        if (z) {
            fq x, y16;
            fq_from_words<8>(x, w);
            fq_from_words<4>(y16, w + 12);
            fq_xyzz_madd(acc, x, y16);
        }
    }
    xyzz o, c;
    fq_xyzz_to_words(o.x.l, o.y.l, o.zz.l, o.zzz.l, acc);
    xyzz_canon(c, o);
    store_xyzz(&out[g], c);
}

void launch_gather_probe(hipStream_t stream, int stride_bytes, const uint4* pts, const uint32_t* idx, void* out,
                         uint32_t nentries) {
    const uint32_t G = nentries / kGatherK;
    const dim3 grid(G / kProbeThreads), block(kProbeThreads);
    if (stride_bytes == 128) hipLaunchKernelGGL(k_gather_sums<8>, grid, block, 0, stream, pts, idx, (xyzz*)out, G);
    else hipLaunchKernelGGL(k_gather_sums<6>, grid, block, 0, stream, pts, idx, (xyzz*)out, G);
}

// lanes G = npairs / K; npairs a multiple of 256 K
void launch_ba_probe(hipStream_t stream, int K, bool batch_affine, const uint4* pts, const uint32_t* idx, fp* pi,
                     void* out, uint32_t npairs) {
    const uint32_t G = npairs / (uint32_t)K;
    const dim3 grid(G / kProbeThreads), block(kProbeThreads);
    if (batch_affine) {
        if (K == 8) hipLaunchKernelGGL(k_ba_pairs<8>, grid, block, 0, stream, pts, idx, pi, (aff*)out, G);
        else if (K == 32) hipLaunchKernelGGL(k_ba_pairs<32>, grid, block, 0, stream, pts, idx, pi, (aff*)out, G);
        else hipLaunchKernelGGL(k_ba_pairs<128>, grid, block, 0, stream, pts, idx, pi, (aff*)out, G);
    } else {
        if (K == 8) hipLaunchKernelGGL(k_madd_pairs<8>, grid, block, 0, stream, pts, idx, (xyzz*)out, G);
        else if (K == 32) hipLaunchKernelGGL(k_madd_pairs<32>, grid, block, 0, stream, pts, idx, (xyzz*)out, G);
        else hipLaunchKernelGGL(k_madd_pairs<128>, grid, block, 0, stream, pts, idx, (xyzz*)out, G);
    }
}

}  // namespace bls
