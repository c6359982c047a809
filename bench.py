#!/usr/bin/env python3
"""Benchmark of the gfx950 G1 MSM backend.

One step = one complete G1 MSM of n = 2^20 points (window digits, sort,
bucket accumulation, bucket reduction, window fold; the 144-byte blst_p1
result on the host) over inputs already resident in HBM.  Workload: P_i =
[k_i]G with k_i uniform in Fr, uniform Fr scalars (nbits = 255), with the
edge-case mix 1% zero / 1% one / 1% r-1 scalars and 0.5% repeated points, all
from seeded xorshift64* streams.

Prints one JSON line.  The roofline object prices the dominant kernel
(bucket accumulation) against the VALU 32-bit integer multiply-add peak of
MI355X; the cpu_baseline object times the oracle (a C port of the reference
path) on the host, single-threaded, on the same inputs, and its compressed
output must equal the GPU's byte for byte.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "temp-cat-blst_amd"))

import numpy as np  # noqa: E402

METRIC = "G1 MSM points/s @2^20 (bit-exact); batched vs plain GPU throughput + p99 latency"
SEED = 0x9E3779B97F4A7C15

# MI355X: 256 CUs x 64 lanes x 2.4 GHz, v_mad_u64_u32 issuing at the full VALU
# rate (one 32x32->64 multiply-add per lane per clock)
PEAK_INT_MAD_PER_S = 256 * 64 * 2.4e9
INT_MAD_PER_FP_MUL = 300   # 12x12 a*b + 12x12 m*p + 12 m_i, Montgomery on 32-bit limbs


def algorithmic_fp_muls(n):
    """W(n) = 6 n ceil(256/c) + 28 ceil(256/c) 2^(c-1), minimised over c in [4, 21].

    6 = one batch-affine bucket addition, 28 = the two XYZZ additions per
    bucket of the running sum.  Returns (W, c*, accumulation share)."""
    best = None
    for c in range(4, 22):
        d = math.ceil(256 / c)
        w = 6 * n * d + 28 * d * 2 ** (c - 1)
        if best is None or w < best[0]:
            best = (w, c, 6 * n * d)
    return best


def load_traffic():
    """HBM bytes per k_accumulate launch, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated)."""
    path = os.path.join(ROOT, "profiles", "msm_2pow20_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch k_accumulate:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(points, scalars, n, gpu_compressed, samples, jac=None, wbits_case=None, g2_case=None):
    """The oracle (C port of blst_p1s_mult_pippenger) on the host cores, and
    what checks sampled batched results; with jac, also the oracle's batch
    normalisation on a 2^18-point sample of it; with wbits_case, the oracle's
    table MSM on the same table and scalars."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import blst_hip as bh

    checked = {}
    for name, items in samples.items():
        ok = True
        for pts, sc, res in items:
            ok = ok and oracle.compress(oracle.msm_pippenger(pts, sc, pts.shape[0], 255, threads=16)) == bh.compress(res)
        checked[name] = ok

    t = time.perf_counter()
    r = oracle.msm_pippenger(points, scalars, n, 255, threads=1)
    dt = time.perf_counter() - t
    cpu_bytes = oracle.compress(r)
    threads = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    r_mt = oracle.msm_pippenger(points, scalars, n, 255, threads=threads)
    dt_mt = time.perf_counter() - t
    assert oracle.compress(r_mt) == cpu_bytes
    aff_rep = None
    if jac is not None:
        m = min(1 << 18, jac.shape[0])
        t = time.perf_counter()
        a = oracle.p1s_to_affine(jac[:m])
        da = time.perf_counter() - t
        aff_rep = {"value": m / da, "unit": "points/s", "cores": 1, "kind": "port",
                   "sample": "oracle points_to_affine over the first %d Jacobian points, 1 thread, %.3f s" % (m, da),
                   "equals_input": bool(np.array_equal(a, points[:m]))}
    wb_rep = None
    if wbits_case is not None:
        table, wb, ssc, gpu_wb = wbits_case
        t = time.perf_counter()
        rw = oracle.mult_wbits(table, wb, ssc, ssc.shape[0], 255)
        dw = time.perf_counter() - t
        wb_rep = {"value": ssc.shape[0] / dw, "unit": "points/s", "cores": 1, "kind": "port",
                  "sample": "oracle msm_mult_table over the 4096-point wbits = 10 table, 1 thread, %.3f s" % dw,
                  "bit_exact_vs_gpu": oracle.compress(rw) == gpu_wb}
    g2_rep = None
    if g2_case is not None:
        gp, gs, gr = g2_case
        t = time.perf_counter()
        w2 = oracle.msm_p2_pippenger(gp, gs, gp.shape[0], 255)
        d2 = time.perf_counter() - t
        g2_rep = {"value": gp.shape[0] / d2, "unit": "points/s", "cores": 1, "kind": "port",
                  "sample": "oracle G2 Pippenger over the first %d bases, 1 thread, %.3f s" % (gp.shape[0], d2),
                  "bit_exact_vs_gpu": oracle.p2_compress(w2) == oracle.p2_compress(gr)}
    return {
        "value": n / dt,
        "unit": "points/s",
        "cores": 1,
        "kind": "port",
        "sample": "one full MSM of the same n=2^20 inputs, oracle Pippenger (reference window %d), 1 thread, %.2f s"
        % (oracle.window_size(n), dt),
        "all_cores": {"value": n / dt_mt, "threads": threads, "seconds": dt_mt},
        "bit_exact_vs_gpu": cpu_bytes == gpu_compressed,
        "compressed": cpu_bytes.hex(),
        "batched_samples_bit_exact": checked,
        "p1s_to_affine": aff_rep,
        "p1s_mult_wbits": wb_rep,
        "p2s_mult_pippenger": g2_rep,
    }


def _same_p2(a, b):
    """Two blst_p2 Jacobian results name the same point: X1 Z2^2 = X2 Z1^2
    and Y1 Z2^3 = Y2 Z1^3 over Fp2, checked on the device-independent limbs
    by Python big integers (Montgomery factors cancel on both sides)."""
    P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB

    def f2(v, k):
        return (sum(int(x) << (64 * i) for i, x in enumerate(v[12 * k:12 * k + 6])),
                sum(int(x) << (64 * i) for i, x in enumerate(v[12 * k + 6:12 * k + 12])))

    def mul(x, y):
        return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)

    ax, ay, az = f2(a, 0), f2(a, 1), f2(a, 2)
    bx, by, bz = f2(b, 0), f2(b, 1), f2(b, 2)
    if az == (0, 0) or bz == (0, 0):
        return az == bz
    az2, bz2 = mul(az, az), mul(bz, bz)
    return mul(ax, bz2) == mul(bx, az2) and mul(ay, mul(bz2, bz)) == mul(by, mul(az2, az))


def entry_points(bh, points, scalars, r, n, reps=5):
    """The other drop-in entry points on the same n-point workload, each
    checked against the plain MSM result or its own input: the blst_fr MSM and
    the plain MSM from host buffers (PCIe-inclusive), and batch normalisation
    of Jacobian representatives of the points, on device memory and host to
    host.  Returns (report, Jacobian points)."""
    fr = bh.synth_fr(scalars)
    # one untimed call of each at the full size first: the first call sizes
    # the device workspaces of its point ranges
    bh.p1s_mult_fr(points, fr, "pippenger")
    t = time.perf_counter()
    for _ in range(reps):
        ok, rf = bh.p1s_mult_fr(points, fr, "pippenger")
    fr_s = (time.perf_counter() - t) / reps
    bh.p1s_mult_pippenger(points, scalars, 255)
    t = time.perf_counter()
    for _ in range(reps):
        rp = bh.p1s_mult_pippenger(points, scalars, 255)
    plain_s = (time.perf_counter() - t) / reps
    fr_same = ok and bh.compress(rf) == bh.compress(r) == bh.compress(rp)

    jac = bh.synth_jacobian(points, SEED ^ 4)
    d_in = bh.DeviceBuffer(jac)
    d_out = bh.DeviceBuffer(np.zeros_like(points))
    try:
        bh.p1s_to_affine_device(d_out.ptr, d_in.ptr, n)
        t = time.perf_counter()
        for _ in range(4 * reps):
            bh.p1s_to_affine_device(d_out.ptr, d_in.ptr, n)
        dev_s = (time.perf_counter() - t) / (4 * reps)
        dev_same = np.array_equal(d_out.read().view(np.uint64).reshape(points.shape), points)
    finally:
        d_in.close()
        d_out.close()
    bh.p1s_to_affine(jac)
    t = time.perf_counter()
    for _ in range(reps):
        aff = bh.p1s_to_affine(jac)
    host_s = (time.perf_counter() - t) / reps
    # a resident 4096-point setup and its wbits = 10 table in HBM
    sn, wb = 4096, 10
    srs = bh.generator_multiples(bh.synth_srs_scalars(sn, SEED ^ 0x20))
    ssc = bh.synth_scalars(sn, SEED ^ 0x21)
    d_srs, d_ssc = bh.DeviceBuffer(srs), bh.DeviceBuffer(ssc)
    d_tab = bh.DeviceAlloc(sn * (1 << (wb - 1)) * 96)
    try:
        t = time.perf_counter()
        bh.wbits_precompute_device(d_tab.ptr, wb, d_srs.ptr, sn)
        pre_s = time.perf_counter() - t
        rw = bh.mult_wbits_device(d_tab.ptr, wb, sn, d_ssc.ptr, 255)
        t = time.perf_counter()
        for _ in range(4 * reps):
            rw = bh.mult_wbits_device(d_tab.ptr, wb, sn, d_ssc.ptr, 255)
        tab_s = (time.perf_counter() - t) / (4 * reps)
        rpl = bh.msm_device(d_srs.ptr, d_ssc.ptr, sn, 255)
        t = time.perf_counter()
        for _ in range(4 * reps):
            rpl = bh.msm_device(d_srs.ptr, d_ssc.ptr, sn, 255)
        pl_s = (time.perf_counter() - t) / (4 * reps)
        table = d_tab.read().view(np.uint64).reshape(-1, 12)
    finally:
        d_srs.close()
        d_ssc.close()
        d_tab.close()
    rh = bh.mult_wbits(table, wb, ssc, 255)
    t = time.perf_counter()
    for _ in range(reps):
        rh = bh.mult_wbits(table, wb, ssc, 255)
    host_tab_s = (time.perf_counter() - t) / reps
    wb_same = bh.compress(rw) == bh.compress(rpl) == bh.compress(rh)
    # G2: 2^16 resident bases through the device entry; a 4096-point sample
    # through the drop-in entry for the oracle's comparison
    gn = 1 << 16
    g2p = bh.synth_p2_points(gn, SEED ^ 0x50)
    g2s = bh.synth_scalars(gn, SEED ^ 0x51)
    d_g2p, d_g2s = bh.DeviceBuffer(g2p), bh.DeviceBuffer(g2s)
    try:
        r2 = bh.p2_msm_device(d_g2p.ptr, d_g2s.ptr, gn, 255)
        t = time.perf_counter()
        for _ in range(4 * reps):
            r2 = bh.p2_msm_device(d_g2p.ptr, d_g2s.ptr, gn, 255)
        g2_s = (time.perf_counter() - t) / (4 * reps)
    finally:
        d_g2p.close()
        d_g2s.close()
    r2h = bh.p2s_mult_pippenger(g2p, g2s, 255)
    g2_sample = (g2p[:4096].copy(), g2s[:4096].copy(), bh.p2s_mult_pippenger(g2p[:4096], g2s[:4096], 255))
    rep = {
        "p2s_mult_pippenger": {
            "workload": "G2 MSM, 2^16 bases [a + i b]G2 and uniform Fr scalars resident in HBM",
            "device_ms": g2_s * 1e3, "device_points_per_s": gn / g2_s,
            "host_entry_equals_device": _same_p2(r2h, r2),
        },
        "p1s_mult_wbits": {
            "workload": "resident 4096-point KZG setup, wbits = 10 table (201 MB) in HBM, one MSM of uniform Fr scalars",
            "precompute_device_ms": pre_s * 1e3,
            "device_ms": tab_s * 1e3, "device_points_per_s": sn / tab_s,
            "plain_pippenger_device_ms": pl_s * 1e3,
            "host_table_ms": host_tab_s * 1e3,
            "equals_plain": wb_same,
        },
        "p1s_mult_pippenger_fr": {
            "workload": "the same n points and scalars as blst_fr (Montgomery form), host buffers in and out",
            "points_per_s": n / fr_s, "ms": fr_s * 1e3,
            "plain_host_points_per_s": n / plain_s, "plain_host_ms": plain_s * 1e3,
            "equals_plain": fr_same,
        },
        "p1s_to_affine": {
            "workload": "n Jacobian representatives (x l^2, y l^3, l) of the same points, seeded l",
            "device_points_per_s": n / dev_s, "device_ms": dev_s * 1e3,
            "host_points_per_s": n / host_s, "host_ms": host_s * 1e3,
            "equals_input": dev_same and np.array_equal(aff, points),
        },
    }
    return rep, jac, (table, wb, ssc, bh.compress(rw)), g2_sample


def ba_probe(npairs=1 << 25):
    """Batch-affine addition against the XYZZ mixed addition of the bucket
    accumulation on the same 2^25 point pairs (test-kit probe kernels), for
    contiguous pairs and for pairs gathered through a random index (as the
    first level of a sorted bucket list would read them); the probe checks
    its sums in tests/test_ba_probe_gpu.py."""
    import blst_hip as bh

    pts = bh.synth_points(2 * npairs, SEED ^ 0xBA)
    idx = np.random.default_rng(SEED & 0xFFFFFFFF).permutation(2 * npairs).astype(np.uint32)
    d_p, d_i = bh.DeviceBuffer(pts), bh.DeviceBuffer(idx)
    rows = []
    try:
        for gather in (False, True):
            for K in (32, 128):
                ns, _, _ = bh.batch_affine_probe(d_p.ptr, npairs, K, d_i.ptr if gather else None, read=False)
                rows.append({"pairs_per_lane": K, "gathered": gather, "batch_affine_ns_per_add": ns[0],
                             "xyzz_ns_per_add": ns[1], "batch_affine_speedup": ns[1] / ns[0]})
    finally:
        d_p.close()
        d_i.close()
    print(json.dumps({"batch_affine_probe": {"pairs": npairs, "rows": rows}}))


def gather_probe(n=1 << 20, windows=16):
    """The accumulation's point gathers at the bench workload's shape (n
    points, each read once per window in a random order: n * windows entries)
    from the 96-byte blst_p1_affine array and from a copy padded to 128-byte
    lines, summed with the accumulation's arithmetic (test-kit probe kernel;
    the sums are checked in tests/test_ba_probe_gpu.py)."""
    import blst_hip as bh

    pts = bh.synth_points(n, SEED ^ 1)
    pad = np.zeros((n, 16), np.uint64)
    pad[:, :12] = pts
    rng = np.random.default_rng(SEED & 0xFFFFFFFF)
    idx = np.concatenate([rng.permutation(n) for _ in range(windows)]).astype(np.uint32)
    d_p, d_q, d_i = bh.DeviceBuffer(pts), bh.DeviceBuffer(pad), bh.DeviceBuffer(idx)
    try:
        ns96, s96 = bh.gather_probe(d_p.ptr, 96, d_i.ptr, idx.size)
        ns128, s128 = bh.gather_probe(d_q.ptr, 128, d_i.ptr, idx.size)
    finally:
        d_p.close()
        d_q.close()
        d_i.close()
    print(json.dumps({"gather_probe": {"points": n, "entries": int(idx.size), "ns_per_entry_96B": ns96,
                                       "ns_per_entry_128B": ns128, "same_sums": bool(np.array_equal(s96, s128))}}))


def skewed_scalars(n=1 << 20, reps=5):
    """Lone MSMs of n device-resident points under scalar mixes that crowd
    the buckets: uniform, 16 distinct values, half the scalars equal, all
    equal, all 1 (every entry in one bucket of one window).  Median and best
    of reps timed calls after one untimed call.  Timing only: results under
    such mixes are checked against the oracle in tests/test_msm_gpu.py
    (heavy-bucket tests) and tests/test_random_sweep_gpu.py."""
    import blst_hip as bh

    pts = bh.synth_points(n, SEED ^ 3)
    uni = bh.synth_scalars(n, SEED ^ 4)
    one_val = np.frombuffer((0x1234567890ABCDEF1122334455667788).to_bytes(32, "little"), np.uint8)
    unit = np.frombuffer((1).to_bytes(32, "little"), np.uint8)
    rng = np.random.default_rng(SEED & 0xFFFFFFFF)
    mixes = {
        "uniform": uni,
        "16 distinct": uni[rng.integers(0, 16, n)],
        "half equal": np.where((rng.random(n) < 0.5)[:, None], one_val[None, :], uni),
        "all equal": np.repeat(one_val[None, :], n, axis=0),
        "all one": np.repeat(unit[None, :], n, axis=0),
    }
    out = {}
    d_p = bh.DeviceBuffer(pts)
    try:
        for name, s in mixes.items():
            d_s = bh.DeviceBuffer(np.ascontiguousarray(s, np.uint8))
            try:
                bh.msm_device(d_p.ptr, d_s.ptr, n, 255)
                ts = []
                for _ in range(reps):
                    t = time.perf_counter()
                    bh.msm_device(d_p.ptr, d_s.ptr, n, 255)
                    ts.append(time.perf_counter() - t)
            finally:
                d_s.close()
            out[name] = {"median_ms": 1e3 * float(np.median(ts)), "best_ms": 1e3 * min(ts)}
            print("%-12s %.2f ms" % (name, out[name]["median_ms"]), file=sys.stderr, flush=True)
    finally:
        d_p.close()
    print(json.dumps({"skewed_scalars": {"points": n, "nbits": 255, "ms": out}}))


def copy_calibration(nbytes, reps=3):
    """Device-to-device copies of a known size: the HIP runtime's blit kernel
    reads and writes nbytes each, so FETCH_SIZE / WRITE_SIZE of these
    dispatches calibrate the counters against true HBM bytes."""
    import ctypes
    import blst_hip as bh

    hip = ctypes.CDLL("libamdhip64.so")
    src = bh.DeviceBuffer(np.zeros(nbytes, np.uint8))
    dst = bh.DeviceBuffer(np.zeros(nbytes, np.uint8))
    for _ in range(reps):
        rc = hip.hipMemcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), ctypes.c_size_t(nbytes), 3)
        assert rc == 0, rc
    hip.hipDeviceSynchronize()
    src.close()
    dst.close()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--pipeline", type=int, default=0,
                    help="window-group pipelines per MSM in the timed steps (0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workloads")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-entry-points", action="store_true",
                    help="skip the blst_fr MSM and batch-normalisation measurements")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="offered request rate (req/s); 0 = --serve-load times the measured capacity")
    ap.add_argument("--serve-load", type=float, default=0.9, help="offered load as a fraction of capacity")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--ba-probe", action="store_true",
                    help="only run the batch-affine against XYZZ addition probe")
    ap.add_argument("--gather-probe", action="store_true",
                    help="only run the point-gather probe at 96- and 128-byte point strides")
    ap.add_argument("--skewed-scalars", action="store_true",
                    help="only time lone 2^20 MSMs under scalar mixes that crowd the buckets")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.ba_probe:
        ba_probe()
        return
    if args.gather_probe:
        gather_probe()
        return
    if args.skewed_scalars:
        skewed_scalars()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the MSM backend runs on a single device "
              "(sharding an MSM across GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    import blst_hip as bh

    if not bh.available():
        raise SystemExit("no gfx950 device visible")
    n = args.n
    t = time.perf_counter()
    points = bh.synth_points(n, SEED ^ 1)
    scalars = bh.synth_scalars(n, SEED ^ 2)
    bh.synth_edge_cases(points, scalars, SEED ^ 3)
    gen_s = time.perf_counter() - t
    print("inputs ready in %.1f s" % gen_s, file=sys.stderr, flush=True)
    dp = bh.DeviceBuffer(points)
    ds = bh.DeviceBuffer(scalars)

    r = None
    bh.set_pipeline(args.pipeline)
    for _ in range(args.warmup):
        r = bh.msm_device(dp.ptr, ds.ptr, n, 255, args.window)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = bh.msm_device(dp.ptr, ds.ptr, n, 255, args.window)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = n * args.steps / elapsed
    print("timed %d steps: %.3f ms/step" % (args.steps, ms_per_step), file=sys.stderr, flush=True)

    # Stage times and the accumulate launch for the roofline come from MSMs
    # run as one pipeline (no window groups overlapping), so each launch has
    # the GPU to itself; HIP events on the pipeline's stream bracket them.
    acc_ms, phases = [], []
    bh.set_pipeline(1)
    for _ in range(max(3, min(args.steps, 10))):
        r1 = bh.msm_device(dp.ptr, ds.ptr, n, 255, args.window)
        ph = bh.last_phase_ms()
        acc_ms.append(ph["accumulate"])
        phases.append(ph)
    bh.set_pipeline(args.pipeline)
    if bh.compress(r1) != bh.compress(r):
        raise SystemExit("one-pipeline and window-group results differ")

    # the multiply-add issue rate measured on this device, reported beside
    # the derived peak (frac stays against the derived figure)
    mad_measured = bh.mad_issue_rate()

    W, cstar, W_acc = algorithmic_fp_muls(n)
    acc_avg_s = float(np.mean(acc_ms)) / 1e3
    achieved = W_acc * INT_MAD_PER_FP_MUL / acc_avg_s
    phase_avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "plain G1 MSM, n=2^20 affine points [k_i]G, uniform Fr scalars (nbits=255) "
                        "with 1% zero / 1% one / 1% r-1 scalars and 0.5% repeated points; inputs resident in HBM",
            "n": n,
            "nbits": 255,
            "window": args.window or None,
            "window_group_pipelines": args.pipeline or "auto",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_accumulate",
            "achieved": achieved / 1e12,
            "peak": PEAK_INT_MAD_PER_S / 1e12,
            "unit": "TOP/s",
            "frac": achieved / PEAK_INT_MAD_PER_S,
            "traffic": load_traffic(),
            "algorithmic_work": "6*n*ceil(256/c*) Fp-mul x 300 int32 mad, c*=%d: %.4g mad per launch"
                                % (cstar, W_acc * INT_MAD_PER_FP_MUL),
            "avg_launch_ms": acc_avg_s * 1e3,
            "peak_measured": mad_measured / 1e12,
            "frac_of_measured": achieved / mad_measured,
            "peak_measured_how": "v_mad_u64_u32 issue probe (blst_hip_mad_issue_rate): 8 independent "
                                 "accumulators per lane, 8 waves per SIMD on every CU",
        },
        "msm_roofline": {
            "W_fp_mul": W,
            "c_star": cstar,
            "fp_mul_per_s": W / (ms_per_step / 1e3),
            "frac": W * INT_MAD_PER_FP_MUL / (ms_per_step / 1e3) / PEAK_INT_MAD_PER_S,
        },
        "phases_ms": phase_avg,
    }
    gpu_bytes = bh.compress(r)
    out["compressed"] = gpu_bytes.hex()
    dp.close()
    ds.close()
    jac = wbits_case = g2_case = None
    if not args.no_entry_points:
        ep, jac, wbits_case, g2_case = entry_points(bh, points, scalars, r, n)
        out["entry_points"] = ep
        print("fr: %.2f ms host, to_affine: %.3f ms device, wbits: %.3f ms device, g2 2^16: %.3f ms device" % (
            ep["p1s_mult_pippenger_fr"]["ms"], ep["p1s_to_affine"]["device_ms"], ep["p1s_mult_wbits"]["device_ms"],
            ep["p2s_mult_pippenger"]["device_ms"]), file=sys.stderr, flush=True)

    import workloads

    samples = {}
    if not args.no_batched:
        kz = workloads.kzg()
        sm = kz.pop("sample")
        samples["kzg"] = [(sm["srs"], sm["coeffs"][i], sm["results"][i]) for i in range(2)]
        print("kzg: %.2f ms batched, %.1fx; comb table %.2f ms" % (
            kz["batched_ms"], kz["speedup"], kz["comb_table"]["batched_ms"] if kz["comb_table"] else float("nan")),
            file=sys.stderr, flush=True)
        un = workloads.uneven()
        sm = un.pop("sample")
        samples["uneven"] = [(sm["points"], sm["scalars"], sm["result"])]
        print("uneven: %.2f ms batched, %.1fx" % (un["batched_ms"], un["speedup"]), file=sys.stderr, flush=True)
        out["batched"] = {"kzg": kz, "uneven": un}
    if not args.no_serving:
        sv = workloads.serving(args.serve_rate or None, args.serve_seconds, args.serve_load)
        print("serving: %.3g pts/s, p99 %.1f ms, %.1fx" % (sv["served_points_per_s"], sv["served_latency_ms"]["p99"],
                                                           sv["speedup"]), file=sys.stderr, flush=True)
        out["serving"] = sv
    if not args.no_cpu:
        cb = cpu_baseline(points, scalars, n, gpu_bytes, samples, jac, wbits_case, g2_case)
        out["cpu_baseline"] = cb
        out["bit_exact"] = cb["bit_exact_vs_gpu"] and all(cb["batched_samples_bit_exact"].values())
        if "batched" in out and out["batched"]["kzg"]["comb_table"]:
            out["bit_exact"] = out["bit_exact"] and out["batched"]["kzg"]["comb_table"]["equals_plain"]
        if "entry_points" in out:
            out["bit_exact"] = (out["bit_exact"] and out["entry_points"]["p1s_mult_pippenger_fr"]["equals_plain"]
                                and out["entry_points"]["p1s_to_affine"]["equals_input"]
                                and out["entry_points"]["p1s_mult_wbits"]["equals_plain"]
                                and out["entry_points"]["p2s_mult_pippenger"]["host_entry_equals_device"]
                                and cb["p2s_mult_pippenger"]["bit_exact_vs_gpu"]
                                and cb["p1s_to_affine"]["equals_input"]
                                and cb["p1s_mult_wbits"]["bit_exact_vs_gpu"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
