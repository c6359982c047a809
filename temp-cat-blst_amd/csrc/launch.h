// Host-side launchers of the device stages; each is defined next to its
// kernel so that the stages compile as separate translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "msm_common.hip.h"

namespace bls {

void launch_k_digits(dim3 grid, dim3 block, hipStream_t stream, const job_t* __restrict__ jobs, const uint32_t* __restrict__ pbase, uint32_t njobs, uint32_t total_points, uint32_t* __restrict__ ent, uint32_t* __restrict__ rank, uint32_t* __restrict__ hist);
void launch_k_fill(dim3 grid, dim3 block, hipStream_t stream, uint32_t* __restrict__ dst, size_t words, uint32_t value);
void launch_k_copy_words(hipStream_t stream, const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t words);
void launch_k_add_words(hipStream_t stream, const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t words,
                        bool first);
void launch_k_scan_tiles(dim3 grid, dim3 block, hipStream_t stream, const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t* __restrict__ tile_sums, uint32_t N);
void launch_k_scan_add(dim3 grid, dim3 block, hipStream_t stream, uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_offs, uint32_t N);
void launch_k_scatter(dim3 grid, dim3 block, hipStream_t stream, const job_t* __restrict__ jobs, const uint32_t* __restrict__ ebase, uint32_t njobs, uint32_t total_entries, const uint32_t* __restrict__ ent, const uint32_t* __restrict__ rank, const uint32_t* __restrict__ off, uint2* __restrict__ sorted);
void launch_k_count(dim3 grid, dim3 block, size_t lds, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, uint32_t* __restrict__ counts);
void launch_k_prefix(dim3 grid, dim3 block, hipStream_t stream, const group_t* __restrict__ groups, const sgroup_t* __restrict__ sg, const uint32_t* __restrict__ gkb, uint32_t ngroups, uint32_t nkeys, uint32_t* __restrict__ counts, uint32_t* __restrict__ hist);
void launch_k_place(dim3 grid, dim3 block, size_t lds, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, const uint32_t* __restrict__ counts, const uint32_t* __restrict__ off, uint2* __restrict__ sorted);
size_t sort_lds_limit();
void launch_k_bin_count(dim3 grid, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, uint32_t* __restrict__ counts, uint32_t ncounts);
void launch_k_bin_place(dim3 grid, hipStream_t stream, const job_t* __restrict__ jobs, const sgroup_t* __restrict__ sg, const uint2* __restrict__ tiles, const uint32_t* __restrict__ pos, uint2* __restrict__ tmp);
void launch_k_bin_fine(dim3 grid, hipStream_t stream, const sgroup_t* __restrict__ sg, const uint32_t* __restrict__ gkb, uint32_t ngroups, const uint32_t* __restrict__ pos, const uint2* __restrict__ tmp, uint2* __restrict__ sorted, uint32_t* __restrict__ cnt, uint32_t* __restrict__ off, uint32_t nkeys, uint32_t ncounts);
void launch_k_fr_to_scalar(hipStream_t stream, uint32_t* __restrict__ io, uint32_t n);
template <class F>
void launch_k_accumulate(dim3 grid, dim3 block, hipStream_t stream, const uint2* __restrict__ sorted, const uint32_t* __restrict__ total, const job_t* __restrict__ jobs, const uint32_t* __restrict__ kbase, uint32_t njobs, xyzz_t<F>* __restrict__ buckets, uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots, uint32_t T, uint32_t nchunks, const uint32_t* __restrict__ prev, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, merge_lists_t lists);
template <class F>
void launch_k_partial_merge(hipStream_t stream, const uint32_t* __restrict__ slot_key, xyzz_t<F>* __restrict__ slots, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t T, uint32_t nchunks, xyzz_t<F>* __restrict__ buckets, merge_lists_t lists);
template <class F>
void launch_k_reduce_segments(dim3 grid, dim3 block, hipStream_t stream, const group_t* __restrict__ groups, const uint32_t* __restrict__ sbase, uint32_t ngroups, uint32_t total_segments, const xyzz_t<F>* __restrict__ buckets, const uint32_t* __restrict__ cnt, xyzz_t<F>* __restrict__ nodes);
template <class F>
void launch_k_reduce_tree(dim3 grid, hipStream_t stream, const uint32_t* __restrict__ blk, const uint4* __restrict__ tn, uint32_t ngroups, const xyzz_t<F>* __restrict__ in, xyzz_t<F>* __restrict__ out, xyzz_t<F>* __restrict__ W, uint32_t last);
void launch_k_quad_check(dim3 grid, dim3 block, hipStream_t stream, const uint4* __restrict__ pts, uint32_t n, uint32_t* __restrict__ bad);
template <class F>
void launch_k_fold(hipStream_t stream, const job_t* __restrict__ jobs, uint32_t njobs, const xyzz_t<F>* __restrict__ W, xyzz_t<F>* __restrict__ out);
void launch_k_fixed_base(dim3 grid, dim3 block, hipStream_t stream, const uint8_t* __restrict__ scalars, uint32_t n, uint32_t nbytes, uint32_t nbits, const uint4* __restrict__ table, xyzz* __restrict__ out);
void launch_k_shift_points(hipStream_t stream, const uint4* __restrict__ in, uint32_t n, uint32_t s, xyzz* __restrict__ out);
void launch_k_comb_sum(hipStream_t stream, const uint4* __restrict__ table, const uint8_t* __restrict__ scalars, uint32_t nbytes, uint32_t nbits, uint32_t vec, uint32_t n, uint32_t wbits, uint32_t D, uint32_t njobs, uint32_t C, uint32_t nch, xyzz* __restrict__ part);
template <class F>
void launch_k_wbits_rows(hipStream_t stream, const uint4* __restrict__ pts, uint32_t row0, uint32_t nrows, uint32_t m, uint32_t L, xyzz_t<F>* __restrict__ out);
template <class F>
void launch_k_wbits_sum(hipStream_t stream, const uint4* __restrict__ table, const uint8_t* __restrict__ scalars, uint32_t nbytes, uint32_t nbits, uint32_t vec, const uint4* __restrict__ list, uint32_t n, uint32_t wbits, uint32_t D, uint32_t C, uint32_t nch, xyzz_t<F>* __restrict__ part);
template <class F>
void launch_k_wbits_tree(hipStream_t stream, const xyzz_t<F>* __restrict__ part, uint32_t nch, uint32_t D, xyzz_t<F>* __restrict__ W);
template <class F>
void launch_k_jac_to_affine(hipStream_t stream, const uint4* __restrict__ in, aff_t<F>* __restrict__ out, uint32_t n, uint32_t T);
template <class F>
void launch_k_to_affine(dim3 grid, dim3 block, hipStream_t stream, const xyzz_t<F>* __restrict__ in, aff_t<F>* __restrict__ out, uint32_t n, uint32_t Q);

}  // namespace bls
