/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * A plain-C CPU restatement of category-blst's G1 multi-scalar
 * multiplication path (blst_p1s_mult_pippenger and what it calls), used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check
 * results.  Nothing in the product library includes, links or calls this.
 *
 * Data layout matches the reference ABI (bindings/blst_types.h:95-152):
 *   Fp element      6 x uint64 little-endian limbs, Montgomery form, R=2^384
 *   affine point    12 x uint64 (x, y); all-zero = point at infinity
 *   Jacobian point  18 x uint64 (X, Y, Z); Z == 0 = point at infinity
 *   scalar          ceil(nbits/8) little-endian bytes
 */
#ifndef BLST_ORACLE_H
#define BLST_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* field (Montgomery) */
void or_fp_mul(uint64_t r[6], const uint64_t a[6], const uint64_t b[6]);
void or_fp_add(uint64_t r[6], const uint64_t a[6], const uint64_t b[6]);
void or_fp_sub(uint64_t r[6], const uint64_t a[6], const uint64_t b[6]);
void or_fp_neg(uint64_t r[6], const uint64_t a[6]);
void or_fp_to_mont(uint64_t r[6], const uint64_t a[6]);
void or_fp_from_mont(uint64_t r[6], const uint64_t a[6]);
void or_fp_inv(uint64_t r[6], const uint64_t a[6]);
uint64_t or_fp_sgn0_pty_mont(const uint64_t a[6]);

/* points */
void or_p1_generator(uint64_t out_affine[12]);
void or_p1_from_affine(uint64_t out[18], const uint64_t a[12]);
void or_p1_add(uint64_t out[18], const uint64_t a[18], const uint64_t b[18]);
void or_p1_double(uint64_t out[18], const uint64_t a[18]);
void or_p1_mult(uint64_t out[18], const uint64_t p[18], const uint8_t *scalar, size_t nbits);
void or_p1_to_affine(uint64_t out[12], const uint64_t p[18]);
int or_p1_is_equal(const uint64_t a[18], const uint64_t b[18]);
int or_p1_affine_on_curve(const uint64_t a[12]);
void or_p1_compress(uint8_t out[48], const uint64_t p[18]);
void or_p1s_to_affine(uint64_t *out_affine, const uint64_t *jacobian, size_t n);
void or_p1s_from_affine_scaled(uint64_t *jacobian, const uint64_t *affine, const uint64_t *lambda, size_t n);

/* Pippenger pieces */
size_t or_pippenger_window_size(size_t npoints);
size_t or_pippenger_scratch_sizeof(size_t npoints);
size_t or_pippenger_window_scratch_sizeof(size_t window);
uint64_t or_booth_digit(const uint8_t *scalar, size_t nbits, size_t bit0, size_t window);

/* blst_fr scalars and the scratch sizes of the blst_fr MSM entry points */
void or_fr_to_scalar(uint8_t out[32], const uint64_t fr[4]);
void or_frs_to_scalars(uint8_t *out, const uint64_t *fr, size_t n);
size_t or_pippenger_fr_scratch_sizeof(size_t npoints);
size_t or_batch_affine_fr_scratch_sizeof(size_t npoints, size_t window, size_t batch);
size_t or_msm_fr_scratch_sizeof(size_t max_points);
size_t or_p2_pippenger_fr_scratch_sizeof(size_t npoints);
size_t or_p2_batch_affine_fr_scratch_sizeof(size_t npoints, size_t window, size_t batch);
size_t or_p2_msm_fr_scratch_sizeof(size_t max_points);

/* fixed-base tables: n rows of 2^(wbits-1) affine multiples, and the MSM
 * over them (contiguous scalars of ceil(nbits/8) bytes) */
size_t or_wbits_precompute_sizeof(size_t wbits, size_t npoints);
size_t or_wbits_scratch_sizeof(size_t npoints);
void or_wbits_precompute(uint64_t *table, size_t wbits, const uint64_t *points, size_t npoints);
void or_mult_wbits(uint64_t ret[18], const uint64_t *table, size_t wbits, size_t npoints,
                   const uint8_t *scalars, size_t nbits);
/* the encoded digit of a table window (lowest bit rem) of one scalar */
uint64_t or_table_booth(const uint8_t *scalar, size_t nbits, size_t wbits, size_t rem);

/* MSM over contiguous arrays (points: n x 12 limbs, scalars: n x ceil(nbits/8)) */
void or_msm_pippenger(uint64_t ret[18], const uint64_t *points, const uint8_t *scalars,
                      size_t npoints, size_t nbits, size_t window);
void or_msm_pippenger_mt(uint64_t ret[18], const uint64_t *points, const uint8_t *scalars,
                         size_t npoints, size_t nbits, size_t window, int threads);
void or_msm_tile(uint64_t ret[18], const uint64_t *points, const uint8_t *scalars,
                 size_t npoints, size_t nbits, size_t bit0, size_t window);
/* njobs independent MSMs on a pool of threads; rets holds njobs x 18 limbs */
void or_msm_batch_mt(uint64_t *rets, const uint64_t *const *points, const uint8_t *const *scalars,
                     const size_t *ns, const size_t *nbits, size_t njobs, int threads);
void or_msm_naive(uint64_t ret[18], const uint64_t *points, const uint8_t *scalars,
                  size_t npoints, size_t nbits);

/* seeded synthetic inputs */
void or_gen_scalars(uint8_t *out32, size_t n, uint64_t seed, int low64_pct);
void or_gen_points(uint64_t *out_affine, size_t n, uint64_t seed, int threads);
void or_gen_srs(uint64_t *out_affine, size_t n, uint64_t seed, int threads);
void or_apply_edge_cases(uint64_t *points, uint8_t *scalars32, size_t n, uint64_t seed);

/* G2 over Fp2 (blst_oracle_g2.c): Fp2 = (c0, c1) of 6 limbs each; affine 24
 * limbs, Jacobian 36 limbs; scalars as for G1 */
void or_fp2_mul(uint64_t r[12], const uint64_t a[12], const uint64_t b[12]);
void or_p2_generator(uint64_t out_affine[24]);
void or_p2_from_affine(uint64_t out[36], const uint64_t a[24]);
void or_p2_add(uint64_t out[36], const uint64_t a[36], const uint64_t b[36]);
void or_p2_double(uint64_t out[36], const uint64_t a[36]);
void or_p2_to_affine(uint64_t out[24], const uint64_t p[36]);
void or_p2_mult(uint64_t out[36], const uint64_t p[36], const uint8_t *scalar, size_t nbits);
void or_p2s_to_affine(uint64_t *out_affine, const uint64_t *jacobian, size_t n);
void or_p2_compress(uint8_t out[96], const uint64_t p[36]);
void or_msm_p2_tile(uint64_t ret[36], const uint64_t *points, const uint8_t *scalars, size_t npoints,
                    size_t nbits, size_t bit0, size_t window);
void or_msm_p2_pippenger(uint64_t ret[36], const uint64_t *points, const uint8_t *scalars, size_t npoints,
                         size_t nbits, size_t window);
void or_gen_p2_points(uint64_t *out_affine, size_t n, uint64_t seed);
/* G2 fixed-base tables: rows of 2^(wbits-1) affine multiples (24 limbs
 * each) and the MSM over them, as or_wbits_precompute / or_mult_wbits */
void or_p2_wbits_precompute(uint64_t *table, size_t wbits, const uint64_t *points, size_t npoints);
void or_p2_mult_wbits(uint64_t ret[36], const uint64_t *table, size_t wbits, size_t npoints,
                      const uint8_t *scalars, size_t nbits);

#ifdef __cplusplus
}
#endif
#endif
