"""Autotuner for the gfx950 MSM backend.

For every MSM size class 2^k it measures, on the GPU it runs on, the window
width c (which is also the radix of the one-pass counting sort: 2^(c-1)
bucket keys per window), the accumulation tile T (sorted bucket entries per
lane) and the reduction tile L (buckets per lane), by coordinate descent
(every window near the starting one with the small tiles T and every
segment L, then the larger tiles), separately for

  single  one MSM per call (latency of the whole pipeline), and
  batch   a job inside a batch of same-size jobs totalling about 2^21 points
          (throughput; the per-batch fixed costs are shared),

and writes the table the library loads at start-up (tuning.txt next to
libblst_hip.so, or the path in BLST_HIP_TUNING).

    python3 temp-cat-blst_amd/autotune.py [--out PATH] [--single 6-20] [--batch 6-20|none]

With --nbits 64 or 128 it measures lone MSMs of scalars that wide (rows
"single64" / "single128", used for 33..64-bit and 65..128-bit scalars);
--append adds the rows to an existing table instead of replacing it.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import blst_hip as bh  # noqa: E402

T_CHOICES = (8, 16, 32, 64, 128)
L_CHOICES = (2, 4, 8, 16)
# tiles tried together with each window: small lone MSMs are latency-bound,
# and their best window depends on the tile (a window fixed at a tile of 32
# first left 2^6 .. 2^14 points about 20% slower than their best)
T_JOINT = (8, 16, 32)
BATCH_POINTS = 1 << 21


def _median_time(fn, reps):
    """median seconds of fn; a shape the library refuses (a batch too large
    for one launch) counts as never best"""
    try:
        fn()
    except RuntimeError:
        return float("inf")
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def _descend(measure, c0):
    """every (window, small tile, segment) combination around the starting
    window, then every tile at the best window and segment"""
    best = None
    for c in range(max(1, c0 - 3), min(20, c0 + 3) + 1):
        for T in T_JOINT:
            for L in L_CHOICES:
                t = measure(c, T, L)
                if best is None or t < best[0]:
                    best = (t, c, T, L)
    _, c, T, L = best
    for T2 in T_CHOICES:
        if T2 not in T_JOINT:
            t = measure(c, T2, L)
            if t < best[0]:
                best = (t, c, T2, L)
    return best


def tune(single_range, batch_range, reps=3, log=print, nbits=255):
    n_max = 1 << max(list(single_range) + list(batch_range) + [21])
    points = bh.synth_points(min(n_max, BATCH_POINTS * 2), 0xA770)
    scalars = bh.synth_scalars(points.shape[0], 0xA771)
    mode = "single" if nbits >= 200 else "single%d" % nbits
    if nbits < 200:
        scalars = np.ascontiguousarray(scalars[:, :(nbits + 7) // 8])
    d_p = bh.DeviceBuffer(points)
    d_s = bh.DeviceBuffer(scalars)
    rows = []
    try:
        for lg in single_range:
            n = 1 << lg
            c0 = bh.tuning_get(0, n, nbits)[0]

            def measure(c, T, L):
                return _median_time(lambda: bh.msm_device(d_p.ptr, d_s.ptr, n, nbits, c, T, L), reps)

            t, c, T, L = _descend(measure, c0)
            rows.append((mode, lg, c, T, L, t))
            log("%s 2^%-2d c=%-2d T=%-2d L=%-2d %.3f ms" % (mode, lg, c, T, L, t * 1e3))
        for lg in batch_range:
            n = 1 << lg
            k = max(1, BATCH_POINTS // n)
            total = min(k * n, points.shape[0])
            k = total // n
            jobs = [(d_p.ptr + 96 * j * n, d_s.ptr + 32 * j * n, n, 255) for j in range(k)]
            c0 = bh.tuning_get(1, n)[0]

            def measure(c, T, L):
                bh.tuning_set(1, lg, c, T, L)
                return _median_time(lambda: bh.msm_batch_device(jobs), reps)

            t, c, T, L = _descend(measure, c0)
            bh.tuning_set(1, lg, c, T, L)
            rows.append(("batch", lg, c, T, L, t))
            log("batch  2^%-2d x%-4d c=%-2d T=%-2d L=%-2d %.3f ms" % (lg, k, c, T, L, t * 1e3))
    finally:
        d_p.close()
        d_s.close()
    return rows


def write_table(rows, path, append=False):
    with open(path, "a" if append else "w") as f:
        if not append:
            f.write("# mode log2n window chunk segment   (measured ms)\n")
        for mode, lg, c, T, L, t in rows:
            f.write("%s %d %d %d %d   # %.4f\n" % (mode, lg, c, T, L, t * 1e3))


def _range(s):
    if s == "none":
        return range(0)
    a, b = s.split("-")
    return range(int(a), int(b) + 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    ap.add_argument("--single", default="6-20")
    ap.add_argument("--batch", default="6-20")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--nbits", type=int, default=255, choices=(64, 128, 255))
    ap.add_argument("--append", action="store_true")
    args = ap.parse_args()
    if not bh.available():
        raise SystemExit("no gfx950 device visible")
    batch = _range(args.batch) if args.nbits == 255 else range(0)
    rows = tune(_range(args.single), batch, args.reps, log=lambda m: print(m, flush=True), nbits=args.nbits)
    write_table(rows, args.out, args.append)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
