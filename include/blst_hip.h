/*
 * blst_hip.h -- C ABI of the MI355X (gfx950) backend for BLS12-381 G1
 * multi-scalar multiplication.
 *
 * The first group of entry points is the drop-in boundary: the same names,
 * argument meaning, scratch sizes and fail-closed behaviour as the scalar
 * backend's C API, so a caller linked against this library instead of the
 * reference's MSM objects needs no source change.  The second group exposes
 * what the device path adds on top (device-resident inputs, batches of MSMs,
 * the serving loop, error reporting); it has no counterpart in the reference.
 *
 * All pointers are plain host pointers unless the name says "device".  No
 * entry point throws; every failure writes the point at infinity (all-zero
 * blst_p1), as the reference's early returns do, and records a message that
 * blst_hip_last_error() returns.
 */
#ifndef BLST_HIP_H
#define BLST_HIP_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef BLST_TYPES_H
/* Layout-identical restatement of bindings/blst_types.h:40-42, 95-131 for
 * callers that do not include the reference headers. */
typedef uint8_t blst_byte;
typedef uint64_t blst_limb_t;
typedef struct { blst_limb_t l[384 / 8 / sizeof(blst_limb_t)]; } blst_fp;
typedef struct { blst_fp x, y, z; } blst_p1;
typedef struct { blst_fp x, y; } blst_p1_affine;
typedef struct { blst_limb_t l[256 / 8 / sizeof(blst_limb_t)]; } blst_fr;
typedef struct { blst_fp fp[2]; } blst_fp2;
typedef struct { blst_fp2 x, y, z; } blst_p2;
typedef struct { blst_fp2 x, y; } blst_p2_affine;
#endif

/* ------------------------------------------------------------------ drop-in */

/* Replaces blst_p1s_mult_pippenger_scratch_sizeof
 * (reference bindings/blst_scalar.h:535, blst/c_api.cpp:1371-1374):
 * 192 * 2^(window-1) bytes for the reference's window heuristic, 0 for no points. */
size_t blst_p1s_mult_pippenger_scratch_sizeof(size_t npoints);

/* Replaces blst_p1s_mult_pippenger_window_scratch_sizeof
 * (bindings/blst_scalar.h:537, blst/c_api.cpp:1376-1380): 0 unless window in [1, 24]. */
size_t blst_p1s_mult_pippenger_window_scratch_sizeof(size_t window);

/* Replaces blst_p1s_mult_pippenger (bindings/blst_scalar.h:539,
 * blst/c_api.cpp:1382-1418).  Same contract: descriptor arrays with later
 * NULL entries meaning contiguous continuation; npoints == 0 or nbits == 0,
 * nbits > 256, NULL heads, or npoints > 1 with NULL scratch -> infinity;
 * npoints == 1 needs no scratch.  The caller's scratch is accepted and not
 * written: bucket memory lives on the GPU. */
void blst_p1s_mult_pippenger(blst_p1 *ret, const blst_p1_affine *const points[],
                             size_t npoints, const blst_byte *const scalars[],
                             size_t nbits, blst_limb_t *scratch);

/* Replaces blst_p1s_mult_pippenger_window (bindings/blst_scalar.h:603,
 * blst/c_api.cpp:1546-1583): explicit window in [1, 24], else infinity. */
void blst_p1s_mult_pippenger_window(blst_p1 *ret, const blst_p1_affine *const points[],
                                    size_t npoints, const blst_byte *const scalars[],
                                    size_t nbits, blst_limb_t *scratch, size_t window);

/* Replaces blst_p1s_tile_pippenger (bindings/blst_scalar.h:605,
 * blst/c_api.cpp:1585-1611, 883-925): the signed Booth digit window starting
 * at bit0, sum_i d_i * P_i, without the window combination. */
void blst_p1s_tile_pippenger(blst_p1 *ret, const blst_p1_affine *const points[],
                             size_t npoints, const blst_byte *const scalars[],
                             size_t nbits, blst_limb_t *scratch, size_t bit0,
                             size_t window);

/* Replaces blst_p1s_mult_pippenger_fr_scratch_sizeof (bindings/blst_scalar.h:541,
 * blst/c_api.cpp:1420-1423, 272-299): 32 bytes per scalar plus, for more than
 * one point, 192 * 2^(window-1) bucket bytes; 0 for no points or an
 * unrepresentable size. */
size_t blst_p1s_mult_pippenger_fr_scratch_sizeof(size_t npoints);

/* Replaces blst_p1s_mult_pippenger_fr (bindings/blst_scalar.h:549,
 * blst/c_api.cpp:1425-1435, 532-580): contiguous points and reduced
 * Montgomery-form blst_fr scalars (converted to canonical scalars on the
 * device).  Returns false, leaving ret unchanged, for a NULL ret, NULL inputs
 * or scratch, or scratch_size below the queried size (and on a device error);
 * npoints == 0 writes infinity and returns true.  The scratch is not written. */
bool blst_p1s_mult_pippenger_fr(blst_p1 *ret, const blst_p1_affine points[], const blst_fr scalars[],
                                size_t npoints, blst_limb_t *scratch, size_t scratch_size);

/* Replaces blst_p1s_mult_fr_scratch_sizeof (bindings/blst_scalar.h:592,
 * blst/c_api.cpp:1502-1506, 758-807): a capacity query, the largest of the
 * Pippenger and batch-affine workspaces the reference's plan may pick for up
 * to max_points. */
size_t blst_p1s_mult_fr_scratch_sizeof(size_t max_points);

/* Replaces blst_p1s_mult_fr (bindings/blst_scalar.h:597,
 * blst/c_api.cpp:1508-1520, 809-851): the same MSM and contract as
 * blst_p1s_mult_pippenger_fr with this function's scratch size; identity
 * points are accepted. */
bool blst_p1s_mult_fr(blst_p1 *ret, const blst_p1_affine points[], const blst_fr scalars[], size_t npoints,
                      blst_limb_t *scratch, size_t scratch_size);

/* Replaces blst_p1s_to_affine (bindings/blst_scalar.h:521,
 * blst/c_api.cpp:1277-1289, multi_scalar.cpp:160-238): Jacobian points in
 * the reference's descriptor convention (a NULL entry after the head
 * continues contiguously) to affine, canonical Montgomery coordinates; points
 * with Z = 0 become the all-zero affine encoding.  npoints == 0 writes
 * nothing; a NULL array or head writes npoints all-zero points.  Coordinates
 * must be reduced (below p), as every blst_p1 the reference returns is.  A
 * device error writes all-zero points and sets blst_hip_last_error(). */
void blst_p1s_to_affine(blst_p1_affine dst[], const blst_p1 *const points[], size_t npoints);

/* Replaces blst_p1s_mult_wbits_precompute_sizeof (bindings/blst_scalar.h:527,
 * blst/c_api.cpp:1310-1316, multi_scalar.h:244-256): 96 * npoints *
 * 2^(wbits-1) bytes; 0 for wbits outside [2, 14] or on overflow. */
size_t blst_p1s_mult_wbits_precompute_sizeof(size_t wbits, size_t npoints);

/* Replaces blst_p1s_mult_wbits_precompute (bindings/blst_scalar.h:529,
 * blst/c_api.cpp:1318-1330, multi_scalar.cpp:268-286, 481-580): row i of the
 * table holds [1]P_i ... [2^(wbits-1)]P_i in affine form (all-zero rows for
 * an all-zero point), built on the device.  Points use the descriptor
 * convention; a packed points array may occupy the table's last npoints
 * entries.  Invalid arguments write nothing. */
void blst_p1s_mult_wbits_precompute(blst_p1_affine table[], size_t wbits, const blst_p1_affine *const points[],
                                    size_t npoints);

/* Replaces blst_p1s_mult_wbits_scratch_sizeof (bindings/blst_scalar.h:531,
 * blst/c_api.cpp:1332-1335, multi_scalar.h:310-315): 144 bytes per point,
 * at most 8192 points' worth. */
size_t blst_p1s_mult_wbits_scratch_sizeof(size_t npoints);

/* Replaces blst_p1s_mult_wbits (bindings/blst_scalar.h:533,
 * blst/c_api.cpp:1337-1369, multi_scalar.cpp:620-707): the MSM over a table
 * from blst_p1s_mult_wbits_precompute with the same wbits.  npoints == 0 or
 * nbits == 0 writes infinity; nbits > 256, wbits outside [2, 14], or a NULL
 * table, scalar head or scratch writes infinity.  The host gathers the signed
 * table entries the scalars' Booth digits select; the device sums them per
 * window.  The scratch is not written. */
void blst_p1s_mult_wbits(blst_p1 *ret, const blst_p1_affine table[], size_t wbits, size_t npoints,
                         const blst_byte *const scalars[], size_t nbits, blst_limb_t *scratch);

/* Replace the G2 Pippenger entry points (bindings/blst_scalar.h:734-740, 768;
 * blst/c_api.cpp:1890-1975, 2053-2080; multi_scalar.cpp:935-943): the same
 * contracts as the G1 functions above over blst_p2 / blst_p2_affine (Fp2
 * coordinates), with 384 * 2^(window-1) scratch bytes (the size of the
 * reference's Pointxyzz<Fp2> buckets). */
size_t blst_p2s_mult_pippenger_scratch_sizeof(size_t npoints);
size_t blst_p2s_mult_pippenger_window_scratch_sizeof(size_t window);
void blst_p2s_mult_pippenger(blst_p2 *ret, const blst_p2_affine *const points[], size_t npoints,
                             const blst_byte *const scalars[], size_t nbits, blst_limb_t *scratch);
void blst_p2s_mult_pippenger_window(blst_p2 *ret, const blst_p2_affine *const points[], size_t npoints,
                                    const blst_byte *const scalars[], size_t nbits, blst_limb_t *scratch,
                                    size_t window);
void blst_p2s_tile_pippenger(blst_p2 *ret, const blst_p2_affine *const points[], size_t npoints,
                             const blst_byte *const scalars[], size_t nbits, blst_limb_t *scratch, size_t bit0,
                             size_t window);

/* Replace blst_p2s_mult_pippenger_fr_scratch_sizeof, blst_p2s_mult_pippenger_fr,
 * blst_p2s_mult_fr_scratch_sizeof and blst_p2s_mult_fr
 * (bindings/blst_scalar.h:742-762; blst/c_api.cpp:1976-1991, 2009-2027): the contracts of
 * the G1 blst_fr entry points over G2, with the reference's G2 scratch sizes
 * (32 bytes per canonical scalar plus 384-byte buckets; for blst_p2s_mult_fr
 * the capacity maximum over its batch-affine plan, which G2 uses from 2^12 to
 * 2^14 points with 200- and 296-byte bucket elements). */
size_t blst_p2s_mult_pippenger_fr_scratch_sizeof(size_t npoints);
bool blst_p2s_mult_pippenger_fr(blst_p2 *ret, const blst_p2_affine points[], const blst_fr scalars[],
                                size_t npoints, blst_limb_t *scratch, size_t scratch_size);
size_t blst_p2s_mult_fr_scratch_sizeof(size_t max_points);
bool blst_p2s_mult_fr(blst_p2 *ret, const blst_p2_affine points[], const blst_fr scalars[], size_t npoints,
                      blst_limb_t *scratch, size_t scratch_size);

/* Replace blst_p2s_to_affine, blst_p2s_mult_wbits_precompute_sizeof,
 * blst_p2s_mult_wbits_precompute, blst_p2s_mult_wbits_scratch_sizeof and
 * blst_p2s_mult_wbits (bindings/blst_scalar.h:722-732; blst/c_api.cpp:
 * 1806-1817, 1829-1888): the contracts of the G1 forms above over G2, with
 * 192-byte table entries (rows of [1]P .. [2^(wbits-1)]P, wbits 2..14) and a
 * scratch of 288 bytes per point up to 4096 points, which is not written. */
void blst_p2s_to_affine(blst_p2_affine dst[], const blst_p2 *const points[], size_t npoints);
size_t blst_p2s_mult_wbits_precompute_sizeof(size_t wbits, size_t npoints);
void blst_p2s_mult_wbits_precompute(blst_p2_affine table[], size_t wbits, const blst_p2_affine *const points[],
                                    size_t npoints);
size_t blst_p2s_mult_wbits_scratch_sizeof(size_t npoints);
void blst_p2s_mult_wbits(blst_p2 *ret, const blst_p2_affine table[], size_t wbits, size_t npoints,
                         const blst_byte *const scalars[], size_t nbits, blst_limb_t *scratch);

/* ------------------------------------------------------- device extensions */

/* Last error message of the calling thread ("" when the last call succeeded). */
const char *blst_hip_last_error(void);

/* 1 when a gfx9 device is present and the kernels load, else 0. */
int blst_hip_available(void);

/* MSM over inputs already resident in device memory: device_points is npoints
 * contiguous blst_p1_affine, device_scalars npoints * ceil(nbits/8) bytes.
 * window 0 selects the tuned window.  Returns 0 on success. */
int blst_hip_msm_device(blst_p1 *ret, const void *device_points, const void *device_scalars,
                        size_t npoints, size_t nbits, size_t window);

/* No counterpart in the reference.  blst_hip_msm_device for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same MSM and result for at most
 * 65536 points (nbits <= 256, tuned window; a wider accumulation tile where
 * the tuned one would not fit), computed entirely in a fixed 96 MiB device
 * array that each GPU the library runs on holds for the life of
 * the process.  The call allocates no device or page-locked memory; the
 * first call on a device takes the array's device address (and sets up the
 * library's device context if no other entry point has).  Every call on a
 * device uses that device's one array, which keeps the last call's contents
 * (the MSM clears what it needs): calls that overlapped on one device would
 * race, so the library serialises them
 * with a lock, and each completes before returning.  Returns 0, or -1 for
 * invalid arguments, more than 65536 points or a device error (*ret is then
 * infinity and blst_hip_last_error() says why). */
int blst_hip_msm_device_noalloc(blst_p1 *ret, const void *device_points, const void *device_scalars,
                                size_t npoints, size_t nbits);

/* Device memory for callers without their own allocator (benchmarks,
 * resident SRS bases): copy `bytes` from host memory into a new device
 * buffer on the current device.  Returns NULL on failure. */
void *blst_hip_device_upload(const void *host, size_t bytes);
int blst_hip_device_write(void *device_dst, const void *host_src, size_t bytes);
int blst_hip_device_read(void *host_dst, const void *device_src, size_t bytes);
void *blst_hip_device_alloc(size_t bytes);   /* uninitialised; NULL on failure */
void blst_hip_device_free(void *device_ptr);

/* One MSM of a batch; points and scalars are device pointers (points
 * 16-byte aligned).  window 0 selects the tuned window. */
typedef struct {
    const void *points;
    const void *scalars;
    size_t npoints;
    size_t nbits;
    size_t window;
} blst_hip_msm_job;

/* A batch of MSMs of any sizes, sharing or not sharing bases, computed by one
 * launch sequence: one digit pass, one sort, one accumulation and one
 * reduction over all jobs, and a device-side window fold per job.
 * rets[k] receives job k (infinity for npoints == 0 or nbits == 0).
 * Returns 0 on success. */
int blst_hip_msm_batch_device(blst_p1 *rets, const blst_hip_msm_job *jobs, size_t njobs);

/* A batch captured once into a HIP graph and replayed: the job shapes and the
 * device input pointers are fixed at creation (the caller rewrites the
 * contents of its scalar / point buffers between runs, e.g. new polynomials
 * over a resident KZG setup).  run() replays the whole launch sequence and
 * writes rets[njobs].  Plans are independent of each other and of the
 * calls above; one plan must not be run from two threads at once. */
typedef struct blst_hip_batch_plan blst_hip_batch_plan;
blst_hip_batch_plan *blst_hip_batch_plan_create(const blst_hip_msm_job *jobs, size_t njobs);
int blst_hip_batch_plan_run(blst_hip_batch_plan *plan, blst_p1 *rets);
void blst_hip_batch_plan_destroy(blst_hip_batch_plan *plan);

/* Serving loop: continuous batching of MSM requests on the current device.
 * hbm_budget_bytes bounds the device memory of admitted requests (inputs
 * plus batch workspace; 0 = 80% of the free memory at creation) and
 * max_batch_points the points of one batch (0 = 2^24).  Requests are admitted
 * in arrival order as memory frees up, copied on a copy stream, and launched
 * in batches queued on one compute stream (at most two on the device while
 * copies, batch set-up and completions overlap them), so requests that arrive
 * while a batch runs join the next one.  Results are identical to a lone MSM.
 * submit, wait and stats may be called from any threads.  destroy finishes
 * the requests in flight first; call it once no thread is inside wait. */
typedef struct blst_hip_server blst_hip_server;
blst_hip_server *blst_hip_server_create(size_t hbm_budget_bytes, size_t max_batch_points);
void blst_hip_server_destroy(blst_hip_server *server);
/* Copies npoints bases to the device once (e.g. a KZG setup); returns a base
 * id >= 0 for later requests, or -1. */
int blst_hip_server_add_bases(blst_hip_server *server, const blst_p1_affine *points, size_t npoints);
/* Queues one MSM over the first npoints resident bases of base_id, or, for
 * base_id < 0, over the npoints host points given.  scalars holds npoints *
 * ceil(nbits/8) host bytes.  The host buffers must stay valid until
 * blst_hip_server_wait returns for the ticket.  Returns a ticket > 0, or -1. */
int64_t blst_hip_server_submit(blst_hip_server *server, int base_id, const blst_p1_affine *points,
                               const blst_byte *scalars, size_t npoints, size_t nbits);
/* Blocks until the ticket completes; writes the result and the latency from
 * submission to completion (ms), and forgets the ticket (a later wait for it
 * returns -1, "unknown ticket"; threads already waiting for it all get the
 * result).  Returns 0 on success. */
int blst_hip_server_wait(blst_hip_server *server, int64_t ticket, blst_p1 *ret, double *latency_ms);
/* [submitted, completed, batches, max batches in flight, peak admitted
 * bytes, HBM budget] */
void blst_hip_server_stats(blst_hip_server *server, uint64_t out[6]);

/* Page-locked host memory (asynchronous copies to the device). */
void *blst_hip_host_alloc(size_t bytes);
void blst_hip_host_free(void *host_ptr);

/* Explicit window and tile shape (chunk = sorted bucket entries per
 * accumulation lane, segment = buckets per reduction lane; 0 = tuned), the
 * hook the autotuner measures with. */
int blst_hip_msm_device_tuned(blst_p1 *ret, const void *device_points, const void *device_scalars,
                              size_t npoints, size_t nbits, size_t window, size_t chunk, size_t segment);

/* Tuning table: (window, chunk, segment) for MSMs of 2^log2n..2^(log2n+1)-1
 * points, for a lone MSM (batch = 0) or a job inside a batch (batch = 1).
 * The library loads the autotuner's table from $BLST_HIP_TUNING or from
 * tuning.txt next to the library; these calls read and override it. */
int blst_hip_tuning_set(int batch, size_t log2n, size_t window, size_t chunk, size_t segment);
void blst_hip_tuning_get(int batch, size_t npoints, size_t nbits, size_t out[3]);

/* Window groups per lone MSM: the windows are split into `groups` pipelines
 * on separate streams (sort, accumulation, merges and reduction of one group
 * overlap those of another).  0 restores the automatic choice (two groups
 * from 2^17 points).  Returns 0, or -1 for groups > 16. */
int blst_hip_set_pipeline(size_t groups);

/* Share of the windows, in percent, given to the last (lowest) window group
 * of a split MSM; the other groups share the rest equally.  The last group's
 * merges, reduction and host fold follow every accumulation, so a smaller
 * share shortens what nothing overlaps.  0 restores equal groups.  Returns 0,
 * or -1 for percent > 50. */
int blst_hip_set_pipeline_share(size_t percent);

/* blst_p1s_to_affine on device memory: npoints Jacobian points (144 bytes
 * each) to affine (96 bytes each), both 16-byte aligned and not overlapping;
 * returns when done, 0 or -1 (npoints below 2^31 - 1). */
/* G2 MSM over inputs resident in device memory (npoints contiguous
 * blst_p2_affine, 16-byte aligned, and npoints * ceil(nbits/8) scalar
 * bytes); window 0 selects the tuned window.  Returns 0 on success. */
int blst_hip_p2_msm_device(blst_p2 *ret, const void *device_points, const void *device_scalars, size_t npoints,
                           size_t nbits, size_t window);

int blst_hip_p1s_to_affine_device(void *device_dst, const void *device_points, size_t npoints);

/* Fixed-base tables resident in device memory (the table of a resident
 * setup never crosses PCIe): blst_p1s_mult_wbits_precompute from device
 * points into a device table of 96 * npoints * 2^(wbits-1) bytes, and
 * blst_p1s_mult_wbits over it with device scalars (npoints contiguous
 * ceil(nbits/8)-byte scalars).  Both return when done, 0 or -1. */
int blst_hip_p1s_mult_wbits_precompute_device(void *device_table, size_t wbits, const void *device_points,
                                              size_t npoints);
int blst_hip_p1s_mult_wbits_device(blst_p1 *ret, const void *device_table, size_t wbits, size_t npoints,
                                   const void *device_scalars, size_t nbits);

/* Comb tables for many MSMs over one resident setup (the KZG shape): window
 * w of the table holds the wbits rows of [2^(wbits w)] P_i, D = 1 +
 * nbits / wbits windows, 96 * D * npoints * 2^(wbits-1) bytes (0 for wbits
 * outside [2, 14], nbits outside [1, 256] or on overflow).  An MSM over it is
 * one sum of D * npoints gathered entries, with no doublings and no window
 * fold.  _build fills a device table from device points; _msm_batch runs
 * njobs MSMs (device scalars, job-major, npoints * ceil(nbits/8) bytes per
 * job) and writes njobs results.  Both return when done, 0 or -1. */
size_t blst_hip_comb_table_bytes(size_t npoints, size_t wbits, size_t nbits);
int blst_hip_comb_table_build(void *device_table, size_t wbits, size_t nbits, const void *device_points,
                              size_t npoints);
int blst_hip_comb_msm_batch(blst_p1 *rets, const void *device_table, size_t wbits, size_t npoints,
                            const void *device_scalars, size_t njobs, size_t nbits);

/* 48-byte ZCash compressed encoding of a G1 Jacobian point, the format of
 * blst_p1_compress (bindings/blst_scalar.h:609, blst/c_api.cpp:1618-1621),
 * computed on the host by this library. */
void blst_hip_p1_compress(blst_byte out[48], const blst_p1 *in);

/* Timing of the phases of the last MSM on this thread, milliseconds:
 * [digits, sort, accumulate, partials, reduce, fold, total]. */
void blst_hip_last_phase_ms(double out[7]);

#ifdef __cplusplus
}
#endif

#endif /* BLST_HIP_H */
