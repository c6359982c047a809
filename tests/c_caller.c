/* A plain C11 caller of the drop-in boundary, built by tests/test_c_caller.py
 * with gcc -std=c11 -pedantic -Wall -Wextra -Werror against
 * include/blst_hip.h and libblst_hip.so, as a C caller of the reference
 * library would be: the same calls, only the library linked differs.
 *
 *   c_caller checks   the scratch sizes of the reference's window heuristic
 *                     (bindings/blst_scalar.h:535-537) and the argument checks
 *                     that return before any device work
 *                     (blst/c_api.cpp:1390-1408); needs no GPU
 *   c_caller msm N    an N-point MSM through blst_p1s_mult_pippenger with
 *                     {base, NULL} descriptors, compressed by the oracle
 *                     (linked in as test infrastructure) and compared with
 *                     the oracle's own MSM of the same inputs; needs a GPU
 *
 * Prints "ok ..." and exits 0, or names the failed check and exits 1. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/blst_hip.h"

#ifdef WITH_ORACLE
#include "../oracle/blst_oracle.h"
#endif

static int failures = 0;

static void check(int ok, const char *what) {
    if (!ok) {
        printf("FAILED %s\n", what);
        failures++;
    }
}

static int is_infinity(const blst_p1 *p) {
    static const blst_p1 zero;
    return memcmp(p, &zero, sizeof zero) == 0;
}

static void run_checks(void) {
    /* reference values of the scratch-size companions */
    check(blst_p1s_mult_pippenger_scratch_sizeof(0) == 0, "scratch size, 0 points");
    check(blst_p1s_mult_pippenger_scratch_sizeof(64) == 3072, "scratch size, 2^6 points");
    check(blst_p1s_mult_pippenger_scratch_sizeof(4096) == 98304, "scratch size, 4096 points");
    check(blst_p1s_mult_pippenger_scratch_sizeof((size_t)1 << 16) == 786432, "scratch size, 2^16 points");
    check(blst_p1s_mult_pippenger_scratch_sizeof((size_t)1 << 18) == 3145728, "scratch size, 2^18 points");
    check(blst_p1s_mult_pippenger_scratch_sizeof((size_t)1 << 20) == 12582912, "scratch size, 2^20 points");
    check(blst_p1s_mult_pippenger_window_scratch_sizeof(0) == 0, "window scratch size, window 0");
    check(blst_p1s_mult_pippenger_window_scratch_sizeof(1) == 192, "window scratch size, window 1");
    check(blst_p1s_mult_pippenger_window_scratch_sizeof(24) == (size_t)192 << 23, "window scratch size, window 24");
    check(blst_p1s_mult_pippenger_window_scratch_sizeof(25) == 0, "window scratch size, window 25");

    blst_p1_affine pts[2];
    blst_byte sc[64];
    memset(pts, 0, sizeof pts);
    memset(sc, 1, sizeof sc);
    const blst_p1_affine *pp[2] = {pts, NULL};
    const blst_byte *ss[2] = {sc, NULL};
    const blst_p1_affine *null_heads[2] = {NULL, NULL};
    blst_limb_t *scratch = malloc(blst_p1s_mult_pippenger_scratch_sizeof(2));
    blst_p1 ret;

    memset(&ret, 7, sizeof ret);
    blst_p1s_mult_pippenger(&ret, pp, 0, ss, 255, scratch);
    check(is_infinity(&ret), "no points -> infinity");
    memset(&ret, 7, sizeof ret);
    blst_p1s_mult_pippenger(&ret, pp, 2, ss, 0, scratch);
    check(is_infinity(&ret), "nbits 0 -> infinity");
    memset(&ret, 7, sizeof ret);
    blst_p1s_mult_pippenger(&ret, pp, 2, ss, 257, scratch);
    check(is_infinity(&ret), "nbits 257 -> infinity");
    memset(&ret, 7, sizeof ret);
    blst_p1s_mult_pippenger(&ret, null_heads, 2, ss, 255, scratch);
    check(is_infinity(&ret), "NULL point head -> infinity");
    memset(&ret, 7, sizeof ret);
    blst_p1s_mult_pippenger(&ret, pp, 2, ss, 255, NULL);
    check(is_infinity(&ret), "two points without scratch -> infinity");
    memset(&ret, 7, sizeof ret);
    blst_p1s_mult_pippenger_window(&ret, pp, 2, ss, 255, scratch, 25);
    check(is_infinity(&ret), "window 25 -> infinity");
    memset(&ret, 7, sizeof ret);
    blst_p1s_tile_pippenger(&ret, pp, 2, ss, 255, scratch, 255, 8);
    check(is_infinity(&ret), "tile past the scalar width -> infinity");
    free(scratch);
}

#ifdef WITH_ORACLE
static int run_msm(size_t n) {
    uint64_t *points = malloc(n * 12 * sizeof(uint64_t));
    uint8_t *scalars = malloc(n * 32 + 32);
    or_gen_points(points, n, 0xCCA11E5ULL, 4);
    or_gen_scalars(scalars, n, 0xCCA11E6ULL, 0);
    const blst_p1_affine *pp[2] = {(const blst_p1_affine *)points, NULL};
    const blst_byte *ss[2] = {scalars, NULL};
    blst_limb_t *scratch = malloc(blst_p1s_mult_pippenger_scratch_sizeof(n));
    blst_p1 ret;
    blst_p1s_mult_pippenger(&ret, pp, n, ss, 255, scratch);
    if (blst_hip_last_error()[0]) printf("error: %s\n", blst_hip_last_error());
    uint64_t want[18];
    uint8_t got_c[48], want_c[48];
    or_msm_pippenger(want, points, scalars, n, 255, 0);
    or_p1_compress(want_c, want);
    or_p1_compress(got_c, (const uint64_t *)&ret);
    check(memcmp(got_c, want_c, 48) == 0, "MSM equals the oracle's");
    free(points);
    free(scalars);
    free(scratch);
    return 0;
}
#endif

int main(int argc, char **argv) {
    if (argc > 1 && strcmp(argv[1], "checks") == 0) {
        run_checks();
#ifdef WITH_ORACLE
    } else if (argc > 2 && strcmp(argv[1], "msm") == 0) {
        run_msm((size_t)strtoull(argv[2], NULL, 10));
#endif
    } else {
        printf("usage: c_caller checks | msm N\n");
        return 2;
    }
    if (failures) return 1;
    printf("ok\n");
    return 0;
}
