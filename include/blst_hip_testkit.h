/*
 * blst_hip_testkit.h -- test and workload-generation hooks of the gfx950 MSM
 * backend, built as libblst_hip_testkit.so beside libblst_hip.so (which it
 * links against).  None of these has a counterpart in the reference and none
 * is part of the drop-in surface: they let tests compare device arithmetic
 * with exact host arithmetic and let benchmarks generate seeded inputs.
 * Errors are reported through blst_hip_last_error() of blst_hip.h.
 */
#ifndef BLST_HIP_TESTKIT_H
#define BLST_HIP_TESTKIT_H

#include "blst_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One device field operation over n elements (48-byte Montgomery limbs, host
 * arrays): op 0 a*b, 1 a^2, 2 a*b + c*d with one reduction, 3 a + b, 4 a - b,
 * 5 2a, 6 -a, 7 -a when bit 0 of b's first limb is set (else a), 8 the
 * canonical representative of a, 9 whether a is zero modulo p (1 or 0 in the
 * first limb), 10 a^-1 (0 for 0).  Inputs must lie in [0, 2p), the device
 * layer's redundant range; outputs lie in [0, 2p) (op 8: [0, p)).  c and d are
 * read only by op 2.  Returns 0 on success, -1 on invalid arguments or a
 * device error. */
int blst_hip_fp_ops(int op, void *out, const void *a, const void *b, const void *c, const void *d, size_t n);

/* One device operation of the 28-bit limb field of the G1 bucket
 * accumulation over n rows of 14 32-bit words (limbs of 28 bits; value
 * sum l_i 2^(28 i), Montgomery radix 2^392): op 0 a*b, 1 a^2, 2 a*b + c*d,
 * 3 a reduced to [0, 2p), 4 a - b + 4p (lazy), 5 4p - a, 6 the X3 step
 * a - (b + 2c) + 6p reduced (a - 2c + 6p when bit 0 of d's first limb is
 * set), 7 whether a is zero modulo p (1 or 0 in the first word), 8, 9, 10
 * the limbs of w 2^8, w 2^4, w where w is the 384-bit number in a's first
 * twelve words, 11 a's value as twelve 32-bit words (then two zero words),
 * 12 4a reduced, 13 a normalized value of any size reduced below 1.03p,
 * 14 the canonical value in [0, p) of a normalized value below 2p.  Inputs
 * must respect each operation's stated limb and value bounds (fq.hip.h).
 * All four inputs are read (pass a for unused ones).
 * Returns 0 on success, -1 on invalid arguments or a device error. */
int blst_hip_fq_ops(int op, uint32_t *out, const uint32_t *a, const uint32_t *b, const uint32_t *c,
                    const uint32_t *d, size_t n);

/* Measured issue rate of v_mad_u64_u32 (32 x 32 -> 64-bit multiply-add with
 * a 64-bit addend, the instruction of every Montgomery product) on the
 * current device: eight independent accumulators per lane, eight waves per
 * SIMD on every CU, best of three timed launches after a warm-up.  Writes
 * multiply-adds per second (all lanes of the device).  Returns 0 or -1. */
int blst_hip_mad_issue_rate(double *mad_per_s);

/* Probe of batch-affine addition against the XYZZ mixed addition of the
 * bucket accumulation, on the same npairs point pairs (device points,
 * 96-byte affine; pair j is points 2j, 2j+1, or device_index[2j],
 * device_index[2j+1] when an index is given): batch-affine with K =
 * pairs_per_lane (8, 32 or 128) pairs per lane and one inversion per
 * 256-lane workgroup, writing the sums as canonical affine points to
 * device_affine_out, and acc = P, acc += Q by madd-2008-s writing canonical
 * XYZZ points to device_xyzz_out.  npairs must be a multiple of 256 K.
 * Writes the best of three timed launches as nanoseconds per pair:
 * [batch-affine, XYZZ].  Returns 0 or -1. */
int blst_hip_batch_affine_probe(const void *device_points, const uint32_t *device_index, size_t npairs,
                                int pairs_per_lane, double ns_per_pair[2], void *device_affine_out,
                                void *device_xyzz_out);

/* Gather probe of the bucket accumulation's memory side: lane g of
 * nentries / 32 lanes sums the affine points device_index[k L + g], k < 32
 * (L lanes), with the accumulation's arithmetic, reading the points at a
 * stride of stride_bytes (96: a blst_p1_affine array; 128: each point padded
 * to one 128-byte line), and writes the canonical XYZZ sums (192 bytes each)
 * to device_xyzz_out.  nentries must be a multiple of 8192.  Writes the best
 * of three timed launches in nanoseconds per entry.  Returns 0 or -1. */
int blst_hip_gather_probe(const void *device_points, size_t stride_bytes, const uint32_t *device_index,
                          size_t nentries, double *ns_per_entry, void *device_xyzz_out);

/* Makes every later device allocation of the product library larger than
 * `bytes` fail as an out-of-memory device would (0 lifts the cap), so tests
 * can drive the fail-closed paths: infinity (or NULL / -1) and a message in
 * blst_hip_last_error(), with later calls unaffected once the cap is lifted. */
void blst_hip_set_alloc_cap(size_t bytes);

/* Makes the staging thread of every later host-buffer MSM split into point
 * ranges (2^19 points or more) fail before it stages point range `range`
 * (-1: never), while the caller's thread waits for that range: the call must
 * come back as infinity with a message in blst_hip_last_error(), and the next
 * call, with the hook off, must be correct. */
void blst_hip_set_stage_failure(int range);

/* Test hook: the signed Booth digits the device computes for window size
 * `window` starting at bit `bit0` (one int32 per scalar, |d| <= 2^(window-1)),
 * written to out[npoints] (host).  Returns 0 on success. */
int blst_hip_booth_digits(int32_t *out, const blst_byte *scalars, size_t npoints,
                          size_t nbits, size_t bit0, size_t window);

/* Test hook: checks the four-lane point additions and doublings of the
 * reduction stages against the one-lane formulas on the affine points
 * (device memory, 96 bytes each, on the curve): for each point P and its
 * successor Q, 2P + 2Q, 2P + 2P in two frames, 2P - 2P, sums with infinity
 * and doublings.  The same cases over the 28-bit limb field, one-lane and
 * four-lane, must give the same elements.  Returns the number of mismatching
 * coordinates, -1 on error. */
int blst_hip_point_ops_check(const void *device_points, size_t npoints);

/* [k_i]G for each scalar (generator multiples, affine, all-zero for k = 0),
 * computed on the device from a comb table of G: the synthetic bases of the
 * benchmark workloads and KZG setup points [tau^i]G.  Returns 0 on success. */
int blst_hip_p1s_mult_generator(blst_p1_affine *out, const blst_byte *scalars, size_t npoints,
                                size_t nbits);

/* Seeded workload generators (xorshift64*): n uniform elements of Fr as
 * 32-byte little-endian scalars (low64_pct percent keep only their low 64
 * bits); the KZG exponents tau^i mod r with tau drawn from the seed; and the
 * edge-case mix (1% zero, 1% one, 1% r-1 scalars, 0.5% repeated points)
 * applied in place at seeded distinct positions; blst_fr (Montgomery form)
 * of canonical scalars below r; Jacobian representatives (x l^2, y l^3, l) of
 * affine points at seeded nonzero l (all-zero points stay all-zero). */
void blst_hip_synth_scalars(blst_byte *out32, size_t n, uint64_t seed, int low64_pct);
void blst_hip_synth_srs_scalars(blst_byte *out32, size_t n, uint64_t seed);
void blst_hip_synth_edge_cases(void *points96, blst_byte *scalars32, size_t n, uint64_t seed);
void blst_hip_synth_fr(uint64_t *out4, const blst_byte *scalars32, size_t n);
void blst_hip_synth_jacobian(void *points144, const void *affine96, size_t n, uint64_t seed);
/* G2 bases [a + i b] G2 (affine, 192 bytes each) for a, b drawn from the seed */
void blst_hip_synth_p2_points(void *points192, size_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* BLST_HIP_TESTKIT_H */
