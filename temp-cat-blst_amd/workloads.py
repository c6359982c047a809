"""Batched and served MSM workloads, each measured against the plain
one-MSM-per-call GPU path on the same inputs.

  kzg       512 MSMs of size 4096 over one resident KZG setup [tau^i]G,
            captured once as a HIP graph and replayed; and the same batch
            over a comb table of the setup (window-shifted fixed-base rows)
  uneven    1000 MSMs, sizes log-uniform in [2^6, 2^18], each with its own
            bases, uniform Fr scalars (10% keep only their low 64 bits), one
            batched launch
  serving   Poisson arrivals of requests in a 70/25/5 mix of 4096 (resident
            setup bases), 2^16 and 2^20 points (fresh bases per request),
            continuous batching vs one call per request in arrival order

Inputs come from the library's seeded generators (xorshift64*); nothing here
computes a group operation on the host.
"""
import math
import threading
import time

import numpy as np

import blst_hip as bh

SEED = 0x9E3779B97F4A7C15


def _rng(seed):
    return np.random.default_rng(seed & 0xFFFFFFFFFFFFFFFF)


def kzg(n_polys=512, n=4096, reps=5, check=True, comb_wbits=13):
    srs = bh.generator_multiples(bh.synth_srs_scalars(n, SEED ^ 0x20))
    coeffs = bh.synth_scalars(n_polys * n, SEED ^ 0x21).reshape(n_polys, n, 32)
    d_srs = bh.DeviceBuffer(srs)
    d_sc = bh.DeviceBuffer(coeffs)
    jobs = [(d_srs.ptr, d_sc.ptr + k * n * 32, n, 255) for k in range(n_polys)]
    plan = bh.BatchPlan(jobs)
    comb = None
    try:
        plan.run()                                    # warm (workspace, code)
        t = time.perf_counter()
        for _ in range(reps):
            got = plan.run()
        batched_s = (time.perf_counter() - t) / reps
        bh.msm_device(jobs[0][0], jobs[0][1], n, 255)
        t = time.perf_counter()
        plain = [bh.msm_device(j[0], j[1], n, 255) for j in jobs]
        plain_s = time.perf_counter() - t
        same = all(bh.compress(got[k]) == bh.compress(plain[k]) for k in range(n_polys)) if check else None
        if comb_wbits:
            # the same batch over a comb table of the resident setup
            tab = bh.DeviceAlloc(bh.comb_table_bytes(n, comb_wbits))
            try:
                t = time.perf_counter()
                bh.comb_table_build(tab.ptr, comb_wbits, d_srs.ptr, n)
                build_s = time.perf_counter() - t
                cg = bh.comb_msm_batch(tab.ptr, comb_wbits, n, d_sc.ptr, n_polys)
                t = time.perf_counter()
                for _ in range(reps):
                    cg = bh.comb_msm_batch(tab.ptr, comb_wbits, n, d_sc.ptr, n_polys)
                comb_s = (time.perf_counter() - t) / reps
            finally:
                tab.close()
            comb = {
                "wbits": comb_wbits,
                "table_bytes": bh.comb_table_bytes(n, comb_wbits),
                "build_ms": build_s * 1e3,
                "batched_ms": comb_s * 1e3,
                "batched_points_per_s": n_polys * n / comb_s,
                "speedup_vs_plain": plain_s / comb_s,
                "equals_plain": all(bh.compress(cg[k]) == bh.compress(plain[k]) for k in range(n_polys)),
            }
    finally:
        plan.close()
        d_srs.close()
        d_sc.close()
    pts = n_polys * n
    return {
        "workload": "%d KZG commitments of degree %d over a resident setup, one HIP graph replay" % (n_polys, n),
        "points": pts,
        "batched_ms": batched_s * 1e3,
        "batched_points_per_s": pts / batched_s,
        "plain_ms": plain_s * 1e3,
        "plain_points_per_s": pts / plain_s,
        "speedup": plain_s / batched_s,
        "batched_equals_plain": same,
        "comb_table": comb,
        "sample": {"srs": srs, "coeffs": coeffs[:2].copy(), "results": got[:2].copy()},
    }


def uneven_sizes(count=1000, lo=6, hi=18, seed=SEED ^ 0x30):
    r = _rng(seed)
    return [int(2 ** (lo + (hi - lo) * u)) for u in r.random(count)]


def uneven(count=1000, check=True):
    sizes = uneven_sizes(count)
    total = sum(sizes)
    points = bh.synth_points(total, SEED ^ 0x31)
    scalars = bh.synth_scalars(total, SEED ^ 0x32, 10)
    d_p = bh.DeviceBuffer(points)
    d_s = bh.DeviceBuffer(scalars)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    jobs = [(d_p.ptr + int(o) * 96, d_s.ptr + int(o) * 32, n, 255) for o, n in zip(offs, sizes)]
    try:
        bh.msm_batch_device(jobs[:4])
        t = time.perf_counter()
        got = bh.msm_batch_device(jobs)
        batched_s = time.perf_counter() - t
        t = time.perf_counter()
        plain = [bh.msm_device(j[0], j[1], j[2], 255) for j in jobs]
        plain_s = time.perf_counter() - t
        same = all(bh.compress(got[k]) == bh.compress(plain[k]) for k in range(count)) if check else None
    finally:
        d_p.close()
        d_s.close()
    small = int(np.argmin(sizes))
    return {
        "workload": "%d MSMs, sizes log-uniform in [2^6, 2^18] (total %d points), distinct bases, "
                    "10%% of scalars 64-bit, one batched launch" % (count, total),
        "points": total,
        "batched_ms": batched_s * 1e3,
        "batched_points_per_s": total / batched_s,
        "plain_ms": plain_s * 1e3,
        "plain_points_per_s": total / plain_s,
        "speedup": plain_s / batched_s,
        "batched_equals_plain": same,
        "sample": {"n": sizes[small], "points": points[offs[small]:offs[small] + sizes[small]].copy(),
                   "scalars": scalars[offs[small]:offs[small] + sizes[small]].copy(), "result": got[small].copy()},
    }


class _ServingData:
    """Resident setup, pools of fresh bases and scalars in page-locked memory."""

    def __init__(self):
        self.srs = bh.generator_multiples(bh.synth_srs_scalars(4096, SEED ^ 0x40))
        self.pinned = []
        self.fresh = {16: [self._pin(bh.synth_points(1 << 16, SEED ^ (0x41 + k))) for k in range(4)],
                      20: [self._pin(bh.synth_points(1 << 20, SEED ^ (0x45 + k))) for k in range(2)]}
        self.scal = {12: [self._pin(bh.synth_scalars(4096, SEED ^ (0x50 + k))) for k in range(32)],
                     16: [self._pin(bh.synth_scalars(1 << 16, SEED ^ (0x80 + k))) for k in range(8)],
                     20: [self._pin(bh.synth_scalars(1 << 20, SEED ^ (0x90 + k))) for k in range(2)]}

    def _pin(self, a):
        p = bh.PinnedArray(a.shape, a.dtype)
        p.array[...] = a
        self.pinned.append(p)
        return p.array

    def close(self):
        for p in self.pinned:
            p.close()


def serving_trace(rate, duration, seed=SEED ^ 0x60):
    """(arrival s, log2 n, fresh-base index, scalar index) of each request."""
    r = _rng(seed)
    t, out = 0.0, []
    while True:
        t += r.exponential(1.0 / rate)
        if t > duration:
            return out
        u = r.random()
        lg = 12 if u < 0.70 else (16 if u < 0.95 else 20)
        out.append((t, lg, int(r.integers(0, 4 if lg == 16 else 2)), int(r.integers(0, {12: 32, 16: 8, 20: 2}[lg]))))


def _pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if x else None


def _serve(data, trace, burst):
    """Runs the trace through a server; burst=True submits everything at once
    (capacity), else at the trace's arrival times.  Returns (seconds, results,
    latencies, stats)."""
    srv = bh.Server()
    res = [None] * len(trace)
    lat = [None] * len(trace)
    try:
        sid = srv.add_bases(data.srs)

        def submit(k):
            _, lg, fi, si = trace[k]
            sc = data.scal[lg][si]
            if lg == 12:
                return srv.submit(sc, 255, base_id=sid, n=4096)
            return srv.submit(sc, 255, points=data.fresh[lg][fi])

        # warm-up before traffic, as a serving process would: a burst of the
        # trace's own first requests (about four full batches) sizes every
        # batch slot's device workspace, which a fresh server allocates on
        # its first full batch
        for t in [submit(k) for k in range(min(900, len(trace)))]:
            srv.wait(t)
        tickets = [None] * len(trace)

        def waiter():
            for k in range(len(trace)):
                while tickets[k] is None:
                    time.sleep(0.0001)
                res[k], lat[k] = srv.wait(tickets[k])

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for k, (at, _, _, _) in enumerate(trace):
            if not burst:
                dt = t0 + at - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            tickets[k] = submit(k)
        th.join()
        secs = time.perf_counter() - t0
        stats = srv.stats()
    finally:
        srv.close()
    return secs, res, lat, stats


def serving(rate=None, duration=2.0, load=0.9, check_every=97, burst_rate=2800.0):
    """Capacity first: after a warm-up burst, a trace of burst_rate req/s
    submitted at once.  Then Poisson arrivals at `rate` req/s, or, without
    one, at `load` times the measured capacity (GPUs and hosts differ by
    several percent, and a fixed rate near one box's capacity overloads a
    slower one); the plain path runs the same Poisson trace one call per
    request."""
    data = _ServingData()
    cap_trace = serving_trace(burst_rate, duration)
    cap_pts = sum(1 << lg for _, lg, _, _ in cap_trace)
    try:
        # a first server in the process pays one-time costs (device memory of
        # the input pool, first launches) that a serving process pays once at
        # start-up: a short burst absorbs them before anything is measured
        _serve(data, cap_trace[:2000], burst=True)
        burst_s, res_burst, _, _ = _serve(data, cap_trace, burst=True)
        capacity = cap_pts / burst_s
        if rate is None:
            rate = load * capacity / (cap_pts / len(cap_trace))
        else:
            load = None
        trace = serving_trace(rate, duration)
        pts_total = sum(1 << lg for _, lg, _, _ in trace)
        served_s, res_served, lat_served, stats = _serve(data, trace, burst=False)

        # plain: one call per request, in arrival order, on the same trace
        d_srs = bh.DeviceBuffer(data.srs)
        d_sc = bh.DeviceBuffer(np.zeros((4096, 32), np.uint8))
        lat_plain, same = [], True

        def plain(lg, fi, si):
            sc = data.scal[lg][si]
            if lg == 12:
                d_sc.write(sc)
                return bh.msm_device(d_srs.ptr, d_sc.ptr, 4096, 255)
            return bh.p1s_mult_pippenger(data.fresh[lg][fi], sc, 255)

        try:
            bh.msm_device(d_srs.ptr, d_sc.ptr, 4096, 255)
            t0 = time.perf_counter()
            for k, (at, lg, fi, si) in enumerate(trace):
                now = time.perf_counter() - t0
                if now < at:
                    time.sleep(at - now)
                r = plain(lg, fi, si)
                lat_plain.append((time.perf_counter() - t0 - at) * 1e3)
                if k % check_every == 0:
                    same = same and bh.compress(r) == bh.compress(res_served[k])
            plain_s = time.perf_counter() - t0
            # the burst's results, sampled, against the plain path
            for k in range(0, len(cap_trace), check_every):
                _, lg, fi, si = cap_trace[k]
                same = same and bh.compress(plain(lg, fi, si)) == bh.compress(res_burst[k])
        finally:
            d_srs.close()
            d_sc.close()
    finally:
        data.close()
    return {
        "workload": "Poisson arrivals at %.0f req/s for %.1f s (%d requests, %d points), mix 70/25/5 of "
                    "4096 (resident setup) / 2^16 / 2^20 (fresh bases)" % (rate, duration, len(trace), pts_total),
        "offered_rate": rate,
        "offered_load_of_capacity": load,
        "offered_points_per_s": pts_total / duration,
        "requests": len(trace),
        "points": pts_total,
        "served_points_per_s": pts_total / served_s,
        "served_latency_ms": {"p50": _pct(lat_served, 50), "p99": _pct(lat_served, 99), "max": max(lat_served)},
        "served_capacity_points_per_s": capacity,
        "capacity_trace": "%d requests (%d points) of a %.0f req/s trace submitted at once" % (
            len(cap_trace), cap_pts, burst_rate),
        "plain_points_per_s": pts_total / plain_s,
        "plain_latency_ms": {"p50": _pct(lat_plain, 50), "p99": _pct(lat_plain, 99), "max": max(lat_plain)},
        "speedup": plain_s / served_s,
        "capacity_speedup": capacity / (pts_total / plain_s),
        "speedup_basis": "speedup: points/s served at the offered load over the plain path's points/s on the same "
                         "trace (the plain path cannot keep up, so its rate is its maximum); capacity_speedup: the "
                         "served form's maximum (the burst above) over the same plain rate",
        "server": stats,
        "served_equals_plain_sampled": same,
    }
