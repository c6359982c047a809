"""Seeded random sweep of the G1 and G2 MSMs against the oracle.

The other GPU tests vary one parameter at a time.  Here each case draws all
of them together from a seeded stream, so the combinations are ones nobody
picked by hand:
* the size (log-uniform, 1 to 20,000 points);
* the scalar width (1 to 256 bits, edge widths more often);
* the window (automatic, or explicit 1 to 24);
* the descriptor convention (contiguous, exact or mixed);
* the scalar mix: uniform, edge values (0, 1, r - 1, r, r + 1, 2^nbits - 1),
  short values, and runs of one repeated scalar, which make heavy buckets;
* the point mix: repeats, negations (P and -P) and all-zero (infinity) entries.

The cases run through the drop-in entry point from host buffers, and through
the device entry points:
* lone MSMs of 2^15 to 2^19 points (window-group pipelines, binned sort);
* batches of uneven jobs in one launch.
A smaller sweep covers G2.  Every result must equal the oracle's 48-byte
(G2: 96-byte) compressed encoding.
"""
import random

import numpy as np
import pytest

import pyref

pytestmark = pytest.mark.gpu
R = pyref.R_ORDER
P = pyref.P


def _nbits(rnd):
    return rnd.choice([1, 2, 7, 8, 9, 31, 32, 33, 63, 64, 65, 127, 128, 129, 200, 253, 254, 255, 255, 255, 256,
                       rnd.randint(1, 256), rnd.randint(1, 256)])


def _scalars(rnd, n, nbits):
    """n scalars of ceil(nbits / 8) bytes (bits above nbits are random: both
    sides ignore them) in one of four mixes."""
    nb = (nbits + 7) // 8
    mix = rnd.choice(["uniform", "edges", "short", "runs"])
    vals = []
    for _ in range(n):
        if mix == "edges" and rnd.random() < 0.5:
            v = rnd.choice([0, 1, R - 1, R, R + 1, (1 << nbits) - 1, 2, 1 << (nbits - 1)])
        elif mix == "short":
            v = rnd.getrandbits(rnd.randint(1, 20))
        elif mix == "runs" and vals and rnd.random() < 0.7:
            v = vals[-1]
        else:
            v = rnd.getrandbits(8 * nb)
        vals.append(v % (1 << (8 * nb)))
    S = np.array([list(v.to_bytes(nb, "little")) for v in vals], dtype=np.uint8).reshape(n, nb)
    return S, mix


def _points(rnd, base, n, limbs=12):
    """n points drawn from base with repeats, negations and infinity."""
    idx = [rnd.randrange(base.shape[0]) for _ in range(n)]
    if n > 4 and rnd.random() < 0.5:                        # a run of one repeated point
        k = rnd.randrange(n)
        for j in range(k, min(n, k + rnd.randint(2, 40))):
            idx[j] = idx[k]
    pts = np.ascontiguousarray(base[idx])
    half = limbs // 2
    for j in range(n):
        u = rnd.random()
        if u < 0.05:
            pts[j] = 0                                      # infinity
        elif u < 0.15:                                      # -P: y -> p - y per Fp coordinate
            for c in range(half // 6):
                lo = half + 6 * c
                y = pyref.limbs_value(pts[j][lo:lo + 6])
                pts[j][lo:lo + 6] = pyref.fp_limbs((P - y) % P)
    return pts


@pytest.fixture(scope="module")
def g1_base(orc):
    return orc.gen_points(4096, 0x5EEB, threads=8)


@pytest.fixture(scope="module")
def g2_base(orc):
    return orc.gen_p2_points(512, 0x2EEB)


def test_sweep_drop_in_host_buffers(hip, orc, g1_base):
    rnd = random.Random(0x5EE1)
    for case in range(150):
        n = int(round(2 ** rnd.uniform(0, np.log2(20000))))
        nbits = _nbits(rnd)
        window = 0 if rnd.random() < 0.7 else rnd.randint(1, 24)
        mode = rnd.choice(["contiguous", "exact", "mixed"])
        k = rnd.randint(1, n)
        pts = _points(rnd, g1_base, n)
        S, mix = _scalars(rnd, n, nbits)
        want = orc.compress(orc.msm_pippenger(pts, S, n, nbits, threads=8 if n > 2000 else 1))
        got = orc.compress(hip.p1s_mult_pippenger(pts, S, nbits, window=window, mode=mode, k=k))
        assert got == want, (case, n, nbits, window, mode, k, mix)


def test_sweep_device_lone_msms(hip, orc, g1_base):
    """Lone device MSMs large enough for two window-group pipelines and the
    binned sort, with the sweep's scalar and point mixes."""
    rnd = random.Random(0x5EE2)
    for case in range(10):
        n = rnd.randint(1 << 15, 1 << 19)
        nbits = rnd.choice([255, 255, 256, 128, 64, _nbits(rnd)])
        pts = _points(rnd, g1_base, n)
        S, mix = _scalars(rnd, n, nbits)
        dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(S)
        try:
            got = hip.compress(hip.msm_device(dp.ptr, ds.ptr, n, nbits))
        finally:
            dp.close()
            ds.close()
        want = orc.compress(orc.msm_pippenger(pts, S, n, nbits, threads=16))
        assert got == want, (case, n, nbits, mix)


def test_sweep_device_batches(hip, orc, g1_base):
    """Batches of 1 to 40 uneven jobs (sizes 0 to 30,000, own scalar width and
    window each, bases shared or not) in one launch sequence."""
    rnd = random.Random(0x5EE3)
    shared = hip.DeviceBuffer(g1_base)
    try:
        for case in range(12):
            jobs, keep, want = [], [], []
            for j in range(rnd.randint(1, 40)):
                n = 0 if rnd.random() < 0.05 else int(round(2 ** rnd.uniform(0, np.log2(30000))))
                nbits = _nbits(rnd)
                window = 0 if rnd.random() < 0.6 else rnd.randint(1, 16)
                S, _ = _scalars(rnd, max(n, 1), nbits)
                if n <= g1_base.shape[0] and rnd.random() < 0.3:
                    pts, dp = g1_base[:max(n, 1)], shared
                else:
                    pts = _points(rnd, g1_base, max(n, 1))
                    dp = hip.DeviceBuffer(pts)
                    keep.append(dp)
                ds = hip.DeviceBuffer(S)
                keep.append(ds)
                jobs.append((dp.ptr, ds.ptr, n, nbits, window))
                want.append((pts[:n], S[:n], n, nbits))
            try:
                got = hip.msm_batch_device(jobs)
            finally:
                for b in keep:
                    b.close()
            res = orc.msm_batch([w for w in want if w[2] > 0], threads=16)
            it = iter(res)
            for j, (_, _, n, nbits) in enumerate(want):
                exp = orc.compress(next(it)) if n else bytes([0xC0]) + bytes(47)
                assert hip.compress(got[j]) == exp, (case, j, n, nbits)
    finally:
        shared.close()


def test_sweep_g2(hip, orc, g2_base):
    rnd = random.Random(0x5EE4)
    for case in range(40):
        n = int(round(2 ** rnd.uniform(0, np.log2(3000))))
        nbits = _nbits(rnd)
        window = 0 if rnd.random() < 0.7 else rnd.randint(1, 16)
        mode = rnd.choice(["contiguous", "exact", "mixed"])
        k = rnd.randint(1, n)
        pts = _points(rnd, g2_base, n, limbs=24)
        S, mix = _scalars(rnd, n, nbits)
        want = orc.p2_compress(orc.msm_p2_pippenger(pts, S, n, nbits))
        got = orc.p2_compress(hip.p2s_mult_pippenger(pts, S, nbits, window=window, mode=mode, k=k))
        assert got == want, (case, n, nbits, window, mode, k, mix)
