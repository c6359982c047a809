"""GPU parity: the gfx950 MSM through the C ABI against the oracle and the
golden fixtures.  Integer work, so every comparison is byte-exact on the
48-byte compressed encoding (the blst_p1 Jacobian representative may differ,
as documented for the boundary)."""
import json
import os
import random

import numpy as np
import pytest

import pyref

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
R = pyref.R_ORDER


def _scal(vals, nb=32):
    return np.array([list(int(v).to_bytes(nb, "little")) for v in vals], dtype=np.uint8).reshape(len(vals), nb)


@pytest.fixture(scope="module")
def base_points(orc):
    return orc.gen_points(4096, 0xB45E, threads=8)


def test_booth_digits_match_oracle(hip, orc):
    rnd = random.Random(21)
    for nbits in (1, 8, 13, 64, 255, 256):
        nb = (nbits + 7) // 8
        ks = [rnd.getrandbits(nbits) for _ in range(300)] + [0, (1 << nbits) - 1]
        S = _scal(ks, nb)
        for c in (1, 2, 5, 8, 11, 16, 17, 24):
            for bit0 in sorted({0, c, 3 * c, nbits - 1, nbits}):
                if bit0 > nbits:
                    continue
                got = hip.booth_digits(S, nbits, bit0, c)
                e = min(nbits - bit0 + 1, c)
                for k, g in zip(ks, got):
                    raw = orc.booth_digit(k.to_bytes(nb, "little"), nbits, bit0, c)
                    mag, sign = raw & ((1 << e) - 1), (raw >> e) & 1
                    assert g == (-mag if sign else mag), (nbits, c, bit0, k)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 33, 100, 513, 4096])
def test_msm_random(hip, orc, base_points, n):
    pts = np.ascontiguousarray(base_points[:n])
    sc = orc.gen_scalars(n, 1000 + n)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert orc.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == want


@pytest.mark.parametrize("nbits", [1, 2, 8, 16, 63, 64, 65, 128, 160, 254, 255, 256])
def test_msm_nbits(hip, orc, base_points, nbits):
    n = 300
    rnd = random.Random(nbits)
    nb = (nbits + 7) // 8
    S = _scal([rnd.getrandbits(8 * nb) for _ in range(n)], nb)   # bits above nbits are ignored
    pts = np.ascontiguousarray(base_points[:n])
    want = orc.compress(orc.msm_pippenger(pts, S, n, nbits))
    assert orc.compress(hip.p1s_mult_pippenger(pts, S, nbits)) == want


@pytest.mark.parametrize("window", [1, 2, 3, 4, 7, 8, 12, 16, 20, 24])
def test_msm_explicit_window(hip, orc, base_points, window):
    n = 200
    pts = np.ascontiguousarray(base_points[:n])
    sc = orc.gen_scalars(n, 7 * window)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert orc.compress(hip.p1s_mult_pippenger(pts, sc, 255, window=window)) == want


@pytest.mark.parametrize("mode", ["contiguous", "exact", "mixed"])
def test_descriptor_conventions(hip, orc, base_points, mode):
    n = 65
    pts = np.ascontiguousarray(base_points[:n])
    sc = orc.gen_scalars(n, 5)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert orc.compress(hip.p1s_mult_pippenger(pts, sc, 255, mode=mode, k=17)) == want


def test_scattered_descriptors(hip, orc, base_points):
    """Exact pointer arrays that are not contiguous (reversed order)."""
    n = 40
    pts = np.ascontiguousarray(base_points[:n])
    sc = orc.gen_scalars(n, 6)
    d = hip.Descriptors(pts, sc, "exact")
    d.pp = d.pp[::-1].copy()
    d.ss = d.ss[::-1].copy()
    ret = np.zeros(18, np.uint64)
    scr = np.zeros(hip.scratch_sizeof(n) // 8, np.uint64)
    hip.load().blst_p1s_mult_pippenger(hip._ptr(ret), hip._ptr(d.pp), n, hip._ptr(d.ss), 255, hip._ptr(scr))
    want = orc.msm_pippenger(np.ascontiguousarray(pts[::-1]), np.ascontiguousarray(sc[::-1]), n, 255)
    assert orc.compress(ret) == orc.compress(want)


def test_single_point_without_scratch(hip, orc, base_points):
    pts = np.ascontiguousarray(base_points[:1])
    for k in (0, 1, 2, R - 1, (1 << 255) - 19):
        sc = _scal([k])
        want = orc.compress(orc.msm_pippenger(pts, sc, 1, 255))
        assert orc.compress(hip.p1s_mult_pippenger(pts, sc, 255, scratch=False)) == want


def test_edge_scalars_and_points(hip, orc, base_points):
    n = 64
    pts = np.array(base_points[:n])
    vals = [int.from_bytes(s.tobytes(), "little") for s in orc.gen_scalars(n, 8)]
    vals[0] = 0
    vals[1] = 1
    vals[2] = R - 1
    vals[3] = R
    vals[4] = (1 << 255) - 1
    pts[10] = pts[11]              # repeated point, same scalar -> doubling in a bucket
    vals[10] = vals[11]
    pts[12] = pts[13]              # repeated point, negated scalar -> cancels in a bucket
    vals[12] = R - vals[13]
    pts[20] = 0                    # all-zero affine = infinity: contributes nothing
    neg = pts[21].copy()           # P and -P with equal scalars
    yv = pyref.limbs_value(neg[6:12])
    neg[6:12] = pyref.fp_limbs((pyref.P - yv) % pyref.P)
    pts[22] = neg
    vals[22] = vals[21]
    sc = _scal(vals)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert orc.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == want
    # naive term-by-term sum agrees too
    assert orc.compress(orc.msm_naive(pts, sc, n, 255)) == want


def test_all_scalars_equal_single_heavy_bucket(hip, orc, base_points):
    n = 4096
    pts = np.ascontiguousarray(base_points[:n])
    sc = _scal([123456789] * n)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert orc.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == want
    sc = _scal([0] * n)
    assert hip.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == bytes([0xC0]) + bytes(47)


def test_heavy_bucket_long_span(hip, orc, base_points):
    # one bucket per window holding all 2^16 entries: its partial sums span
    # far more chunks than one merge workgroup has quads
    n = 1 << 16
    pts = np.ascontiguousarray(np.tile(base_points, (n // base_points.shape[0], 1)))
    sc = _scal([0x1234567890ABCDEF1122334455667788] * n)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert orc.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == want


@pytest.mark.parametrize("k", [1, 0x1234567890ABCDEF1122334455667788])
def test_heavy_bucket_merged_in_parts(hip, orc, base_points, k):
    # 2^18 equal scalars: every window's one bucket holds all entries, 8192
    # tile pieces, merged by several workgroups in parts whose sums the last
    # one to finish adds (k = 1: a single bucket in all)
    n = 1 << 18
    pts = np.ascontiguousarray(np.tile(base_points, (n // base_points.shape[0], 1)))
    sc = _scal([k] * n)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        for _ in range(2):   # the part counters start from zero on every call
            assert hip.compress(hip.msm_device(dp.ptr, ds.ptr, n, 255)) == want
    finally:
        dp.close()
        ds.close()


def test_window_group_pipelines(hip, orc, base_points):
    # the windows split over 1..4 concurrent pipelines give one result
    n = 1 << 17
    pts = np.ascontiguousarray(np.tile(base_points, (n // base_points.shape[0], 1)))
    sc = orc.gen_scalars(n, 0x717E)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        for groups in (1, 2, 3, 4, 0):
            hip.set_pipeline(groups)
            assert hip.compress(hip.msm_device(dp.ptr, ds.ptr, n, 255)) == want, groups
        # unequal groups: the last one with 1, 4 and 7 of the 16 windows
        for groups, share in ((2, 6), (2, 25), (2, 44), (3, 25)):
            hip.set_pipeline(groups)
            hip.set_pipeline_share(share)
            assert hip.compress(hip.msm_device(dp.ptr, ds.ptr, n, 255)) == want, (groups, share)
    finally:
        hip.set_pipeline(0)
        hip.set_pipeline_share(0)
        dp.close()
        ds.close()


def test_quad_point_ops_match_one_lane(hip, base_points):
    pts = np.ascontiguousarray(base_points[:1024])
    d = hip.DeviceBuffer(pts)
    try:
        assert hip.point_ops_check(d.ptr, pts.shape[0]) == 0
    finally:
        d.close()


def test_scalars_beyond_r_256_bits(hip, orc):
    A = np.array([pyref.affine_to_limbs(pyref.G)] * 2, dtype=np.uint64)
    S = np.full((2, 32), 0xFF, np.uint8)
    want = pyref.compress(pyref.mul(2 * ((1 << 256) - 1) % R, pyref.G))
    assert orc.compress(hip.p1s_mult_pippenger(A, S, 256)) == want


def test_fuzz_corpus(hip):
    cases = json.load(open(os.path.join(GOLD, "fuzz_msm.json")))
    for c in cases:
        n = c["n"]
        A = np.array([pyref.affine_to_limbs(pyref.mul(m, pyref.G)) for m in c["point_multipliers"]],
                     dtype=np.uint64)
        S = np.array(c["scalars"], dtype=np.uint8).reshape(n, 1)
        assert hip.compress(hip.p1s_mult_pippenger(A, S, 8)).hex() == c["expected"]
        assert hip.compress(hip.p1s_mult_pippenger(A, S, 8, window=8)).hex() == c["expected"]
        T = np.array(c["tile_scalars"], dtype=np.uint8).reshape(n, 1)
        assert hip.compress(hip.p1s_tile_pippenger(A, T, 8, 0, 8)).hex() == c["tile_expected"]


def test_tile_matches_oracle(hip, orc, base_points):
    n = 500
    pts = np.ascontiguousarray(base_points[:n])
    sc = orc.gen_scalars(n, 31)
    for bit0, w in ((0, 8), (3, 5), (100, 13), (250, 10), (254, 4)):
        want = orc.compress(orc.msm_tile(pts, sc, n, 255, bit0, w))
        assert orc.compress(hip.p1s_tile_pippenger(pts, sc, 255, bit0, w)) == want


def test_hash_to_curve_points(hip):
    pts = json.load(open(os.path.join(GOLD, "h2c_g1.json")))
    aff = [(int(v["x"], 16), int(v["y"], 16)) for v in pts]
    A = np.array([pyref.affine_to_limbs(p) for p in aff], dtype=np.uint64)
    ks = [3 ** i + 7 for i in range(len(aff))]
    want = pyref.compress(pyref.msm(aff, ks))
    assert hip.compress(hip.p1s_mult_pippenger(A, _scal(ks), 255)) == want


def test_generator_multiples_kernel(hip, orc):
    keys = orc.gen_scalars(3000, 4242)
    keys[5] = 0
    keys[6] = 0xFF
    got = hip.generator_multiples(keys)
    want = orc.gen_points(3000, 4242, threads=8)
    want[5] = 0
    g = orc.from_affine(orc.generator_affine())
    want[6] = orc.to_affine(orc.p1_mult(g, bytes(keys[6]), 255))   # bit 255 is outside nbits
    assert np.array_equal(got, want)


def test_config0_n4096(hip, orc):
    """The reference's CPU-runnable case: n=4096 uniform, nbits 255."""
    n = 4096
    pts = hip.synth_points(n, 0x9E3779B97F4A7C15 ^ 0x10)
    sc = hip.synth_scalars(n, 0x9E3779B97F4A7C15 ^ 0x11)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert hip.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == want


def test_config1_n2pow20_with_edge_cases(hip, orc):
    """The plain 2^20 workload with its edge cases, bit-exact against the oracle."""
    n = 1 << 20
    seed = 0x9E3779B97F4A7C15
    pts = hip.synth_points(n, seed ^ 1)
    sc = hip.synth_scalars(n, seed ^ 2)
    hip.synth_edge_cases(pts, sc, seed ^ 3)
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        got = hip.msm_device(dp.ptr, ds.ptr, n, 255)
    finally:
        dp.close()
        ds.close()
    want = orc.msm_pippenger(pts, sc, n, 255, threads=16)
    assert hip.compress(got) == orc.compress(want)
    # the host-buffer entry splits this size into point ranges copied while
    # the previous range runs
    assert hip.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == orc.compress(want)


def test_2pow21_point_ranges_with_edge_cases(hip, orc):
    """At 2^21 points both the device-memory entry and the host-buffer entry
    run the MSM in point ranges sharing one set of buckets; both equal the
    oracle bit for bit."""
    n = 1 << 21
    seed = 0x2F21
    pts = hip.synth_points(n, seed ^ 1)
    sc = hip.synth_scalars(n, seed ^ 2)
    hip.synth_edge_cases(pts, sc, seed ^ 3)
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        got = hip.msm_device(dp.ptr, ds.ptr, n, 255)
    finally:
        dp.close()
        ds.close()
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255, threads=16))
    assert hip.compress(got) == want
    assert hip.compress(hip.p1s_mult_pippenger(pts, sc, 255)) == want


@pytest.mark.parametrize("nbits,window", [(64, 0), (128, 0), (255, 13)])
def test_device_point_ranges_ragged(hip, orc, nbits, window):
    """Device-memory MSMs above 2^21 points run in point ranges: a ragged size
    with short scalars (their own tuned windows) or a caller's window, with
    zero scalars and infinity points in several ranges, against the oracle."""
    n = (1 << 21) + 4099
    nb = (nbits + 7) // 8
    pts = hip.synth_points(n, 0x7A + nbits)
    sc = np.ascontiguousarray(hip.synth_scalars(n, 0x7B + nbits)[:, :nb])
    for i in (0, 1, n // 3, n // 2 + 5, n - 1):
        sc[i] = 0
        pts[i + 7 if i + 7 < n else i - 7] = 0
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        got = hip.msm_device(dp.ptr, ds.ptr, n, nbits, window)
    finally:
        dp.close()
        ds.close()
    want = orc.msm_pippenger(pts, sc, n, nbits, threads=16)
    assert hip.compress(got) == orc.compress(want)


@pytest.mark.parametrize("n,nbits,window", [((1 << 19) + 77, 255, 0), ((1 << 19) + 1000, 13, 0),
                                            (600001, 64, 12), ((1 << 19) + 5, 256, 16)])
def test_host_buffers_split_ranges(hip, n, nbits, window):
    """Host-buffer MSMs large enough to be split into point ranges (ragged
    sizes, short scalars, a caller's window) equal the unsplit MSM over the
    same inputs in device memory."""
    nb = (nbits + 7) // 8
    pts = hip.synth_points(n, 0x5A + n)
    sc = np.ascontiguousarray(hip.synth_scalars(n, 0x5B + n)[:, :nb])
    sc[7] = 0
    sc[n - 3] = 0xFF
    pts[11] = 0
    pts[n - 1] = 0
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        want = hip.msm_device(dp.ptr, ds.ptr, n, nbits, window)
    finally:
        dp.close()
        ds.close()
    assert hip.compress(hip.p1s_mult_pippenger(pts, sc, nbits, window=window)) == hip.compress(want)


@pytest.mark.parametrize("mode,k", [("exact", 1), ("mixed", 37)])
def test_host_buffers_split_ranges_descriptors(hip, mode, k):
    """A split-size MSM whose points and scalars come as per-element pointer
    arrays (the reference's descriptor forms) equals the device-memory MSM."""
    n = (1 << 19) + 333
    pts = hip.synth_points(n, 0x6A)
    sc = hip.synth_scalars(n, 0x6B)
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        want = hip.msm_device(dp.ptr, ds.ptr, n, 255)
    finally:
        dp.close()
        ds.close()
    got = hip.p1s_mult_pippenger(pts, sc, 255, mode=mode, k=k)
    assert hip.compress(got) == hip.compress(want)


def test_fr_split_ranges_ragged(hip, orc):
    """The blst_fr entry at a ragged split size (its scalars converted on the
    device range by range) equals the byte-scalar entry on the oracle's
    conversion of the same values."""
    n = (1 << 19) + 4097
    pts = hip.synth_points(n, 0x6C)
    rnd = np.random.default_rng(0x6D)
    fr = rnd.integers(0, 1 << 63, size=(n, 4), dtype=np.uint64)
    fr[:, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)
    fr[::1000] = 0
    sc = orc.fr_to_scalars(fr)
    ok, got = hip.p1s_mult_fr(pts, fr, "pippenger")
    assert ok
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        want = hip.msm_device(dp.ptr, ds.ptr, n, 255)
    finally:
        dp.close()
        ds.close()
    assert hip.compress(got) == hip.compress(want)


def test_linearity_at_2pow18(hip, orc):
    """MSM(k) + MSM(k') == MSM(k + k' mod r): a size-independent property."""
    n = 1 << 18
    pts = hip.synth_points(n, 99)
    a = hip.synth_scalars(n, 100)
    b = hip.synth_scalars(n, 101)
    ai = np.frombuffer(a.tobytes(), dtype="<u8").reshape(n, 4)
    bi = np.frombuffer(b.tobytes(), dtype="<u8").reshape(n, 4)
    s = np.zeros((n, 32), np.uint8)
    for i in range(0, n, 4096):   # big-int sum mod r in blocks
        for j in range(i, min(n, i + 4096)):
            v = (int.from_bytes(a[j].tobytes(), "little") + int.from_bytes(b[j].tobytes(), "little")) % R
            s[j] = np.frombuffer(v.to_bytes(32, "little"), np.uint8)
    del ai, bi
    ra = hip.p1s_mult_pippenger(pts, a, 255)
    rb = hip.p1s_mult_pippenger(pts, b, 255)
    rs = hip.p1s_mult_pippenger(pts, s, 255)
    assert orc.compress(orc.p1_add(ra, rb)) == orc.compress(rs)


def test_batch_uneven_distinct_and_shared_bases(hip, orc, base_points):
    """One launch sequence over MSMs of different sizes, windows and nbits,
    some sharing their bases; every result equals its own oracle MSM."""
    rnd = random.Random(77)
    sizes = [1, 2, 3, 64, 100, 257, 1000, 4096, 4096, 50, 0, 7]
    nbits_l = [255, 255, 8, 255, 64, 255, 255, 255, 255, 17, 255, 256]
    windows = [0, 0, 0, 5, 0, 3, 0, 0, 11, 0, 0, 2]
    shared = hip.DeviceBuffer(base_points)
    keep, jobs, want = [shared], [], []
    for k, (n, nb, w) in enumerate(zip(sizes, nbits_l, windows)):
        nbytes = (nb + 7) // 8
        sc = np.array([[rnd.getrandbits(8) for _ in range(nbytes)] for _ in range(max(n, 1))], np.uint8)
        if k in (7, 8):            # two MSMs over the shared, resident bases
            pts = np.ascontiguousarray(base_points[:n])
            dp = shared
        else:
            pts = orc.gen_points(max(n, 1), 500 + k, threads=4)
            dp = hip.DeviceBuffer(pts)
            keep.append(dp)
        ds = hip.DeviceBuffer(sc)
        keep.append(ds)
        jobs.append((dp.ptr, ds.ptr, n, nb, w))
        want.append(orc.compress(orc.msm_pippenger(pts[:n], sc[:n], n, nb)) if n else bytes([0xC0]) + bytes(47))
    try:
        got = hip.msm_batch_device(jobs)
    finally:
        for b in keep:
            b.close()
    for k in range(len(jobs)):
        assert hip.compress(got[k]) == want[k], k


def test_batch_binned_sort_mixed_windows(hip, orc, base_points):
    """A batch whose windows all have 2^6 .. 2^15 buckets with most entries in
    windows of 2^12 or more takes the binned two-pass sort: jobs of window
    widths 7 .. 14 bits, a short scalar width, repeated and edge scalars;
    every result equals its own oracle MSM."""
    rnd = random.Random(0xB1)
    r = pyref.R_ORDER
    specs = [(1 << 15, 255, 13), (3000, 255, 9), (5000, 64, 7), (1 << 14, 255, 14), (100, 255, 7),
             (4096, 255, 13)]
    keep, jobs, want = [], [], []
    for k, (n, nb, w) in enumerate(specs):
        nbytes = (nb + 7) // 8
        vals = [rnd.getrandbits(nb) for _ in range(n)]
        for i in range(0, n, 97):
            vals[i] = rnd.choice([0, 1, r - 1 if nb == 255 else 1, vals[0]])
        sc = _scal(vals, nbytes)
        pts = np.ascontiguousarray(np.tile(base_points, (n // 4096 + 1, 1))[:n]) if k % 2 else \
            orc.gen_points(n, 900 + k, threads=8)
        dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
        keep += [dp, ds]
        jobs.append((dp.ptr, ds.ptr, n, nb, w))
        want.append(orc.compress(orc.msm_pippenger(pts, sc, n, nb, threads=8)))
    try:
        got = hip.msm_batch_device(jobs)
    finally:
        for b in keep:
            b.close()
    for k in range(len(jobs)):
        assert hip.compress(got[k]) == want[k], k


def test_graph_plan_replays(hip, orc, base_points):
    """A captured batch replayed several times with new scalars written into
    the same device buffers between replays, with eager work in between;
    every replay equals the eager batch and the oracle."""
    srs = hip.DeviceBuffer(base_points)
    sizes = [4096, 4096, 1000, 33]
    scal = [hip.DeviceBuffer(np.zeros((n, 32), np.uint8)) for n in sizes]
    jobs = [(srs.ptr, scal[k].ptr, n, 255) for k, n in enumerate(sizes)]
    plan = hip.BatchPlan(jobs)
    try:
        for rep in range(3):
            host_sc = [orc.gen_scalars(n, 900 + 10 * rep + k) for k, n in enumerate(sizes)]
            for b, h in zip(scal, host_sc):
                b.write(h)
            got = plan.run()
            eager = hip.msm_batch_device(jobs)
            for k, n in enumerate(sizes):
                want = orc.compress(orc.msm_pippenger(np.ascontiguousarray(base_points[:n]), host_sc[k], n, 255))
                assert hip.compress(got[k]) == want, (rep, k)
                assert hip.compress(eager[k]) == want, (rep, k)
    finally:
        plan.close()
        srs.close()
        for b in scal:
            b.close()


def test_serving_loop_results_match_alone(hip, orc, base_points):
    """Requests of mixed sizes over resident and fresh bases, submitted while
    earlier ones run; every result equals the same MSM computed alone and
    the oracle.  A small HBM budget forces admission to wait for memory."""
    srv = hip.Server(hbm_budget=96 << 20, max_batch_points=1 << 16)
    try:
        sid = srv.add_bases(base_points)
        rnd = random.Random(5)
        reqs = []
        for k in range(40):
            kind = rnd.random()
            if kind < 0.6:
                n = rnd.choice([1, 100, 4096])
                sc = orc.gen_scalars(n, 3000 + k)
                reqs.append((n, sc, None, srv.submit(sc, 255, base_id=sid, n=n)))
            else:
                n = rnd.choice([2, 777, 20000])
                pts = orc.gen_points(n, 4000 + k, threads=8)
                sc = orc.gen_scalars(n, 5000 + k)
                reqs.append((n, sc, pts, srv.submit(sc, 255, points=pts)))
        lat = []
        for n, sc, pts, t in reqs:
            got, ms = srv.wait(t)
            lat.append(ms)
            P = np.ascontiguousarray(base_points[:n]) if pts is None else pts
            alone = hip.p1s_mult_pippenger(P, sc, 255)
            want = orc.compress(orc.msm_pippenger(P, sc, n, 255))
            assert hip.compress(got) == hip.compress(alone) == want, (n, t)
        st = srv.stats()
        assert st["completed"] == 40 and st["submitted"] == 40
        assert st["peak_bytes"] <= st["budget"] + (4096 * 96)
        assert st["batches"] < 40                      # requests were batched together
    finally:
        srv.close()


# ---------------------------------------------------------------- blst_fr entry points

FR_ENTRIES = ["pippenger", "msm"]


@pytest.mark.parametrize("entry", FR_ENTRIES)
def test_fr_known_answer_ten_g(hip, orc, entry):
    """Scalars 1..4 in Montgomery form on four copies of G give [10]G, and a
    scratch one byte short is refused with ret untouched
    (BindingHeader_tests.cpp:642-690)."""
    g = orc.generator_affine()
    pts = np.ascontiguousarray(np.tile(g, (4, 1)))
    fr = orc.fr_from_values([1, 2, 3, 4])
    ok, ret = hip.p1s_mult_fr(pts, fr, entry)
    assert ok
    assert hip.compress(ret) == pyref.compress(pyref.mul(10, pyref.G))
    short = np.full(18, 7, np.uint64)
    ok, out = hip.p1s_mult_fr(pts, fr, entry, scratch_size=hip.fr_scratch_sizeof(4, entry) - 1, ret=short)
    assert not ok and (out == 7).all()


@pytest.mark.parametrize("entry", FR_ENTRIES)
@pytest.mark.parametrize("n", [1, 2, 3, 100, 4096, 6000])
def test_fr_random_matches_oracle(hip, orc, entry, n):
    """Random Montgomery-form scalars with edge values (0, one, r - 1 and
    unreduced 256-bit words) against the oracle's conversion and MSM."""
    rnd = random.Random(n * 7 + len(entry))
    pts = orc.gen_points(n, 0xF00 + n, threads=8)
    vals = [rnd.randrange(R) for _ in range(n)]
    for k, v in enumerate([0, 1, R - 1, 2]):
        if k < n:
            vals[k] = v
    fr = orc.fr_from_values(vals)
    if n > 8:
        fr[5] = np.uint64(0xFFFFFFFFFFFFFFFF)       # 2^256 - 1, above r
        fr[6] = orc.limbs(R, 4)                       # r itself
        fr[7] = orc.limbs(R + 12345, 4)
    sc = orc.fr_to_scalars(fr)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    ok, ret = hip.p1s_mult_fr(pts, fr, entry)
    assert ok and hip.compress(ret) == want


def test_fr_identity_point_and_capacity_scratch(hip, orc):
    """An identity point among 2^12 (which sends the reference's selected
    algorithm back to Pippenger) and a capacity-sized scratch for a smaller
    input, as BindingHeader_tests.cpp:1205-1242 does."""
    cap, n = 1 << 12, 1 << 11
    pts = orc.gen_points(cap, 0xCAFE, threads=8)
    rnd = random.Random(99)
    fr = orc.fr_from_values([rnd.randrange(R) for _ in range(cap)])
    big = hip.fr_scratch_sizeof(cap, "msm")
    ok_p, want_small = hip.p1s_mult_fr(pts[:n], fr[:n], "pippenger")
    ok_m, got_small = hip.p1s_mult_fr(pts[:n], fr[:n], "msm", scratch_size=big)
    assert ok_p and ok_m and hip.compress(got_small) == hip.compress(want_small)
    assert hip.compress(want_small) == orc.compress(orc.msm_pippenger(pts[:n], orc.fr_to_scalars(fr[:n]), n, 255))
    pts[17] = 0
    ok, got = hip.p1s_mult_fr(pts, fr, "msm", scratch_size=big)
    assert ok
    assert hip.compress(got) == orc.compress(orc.msm_pippenger(pts, orc.fr_to_scalars(fr), cap, 255))
    short = np.full(18, 7, np.uint64)
    ok, out = hip.p1s_mult_fr(pts, fr, "msm", scratch_size=big - 1, ret=short)
    assert not ok and (out == 7).all()


def test_fr_2pow20_equals_canonical_scalars(hip, orc):
    """At 2^20 points the blst_fr entry equals the byte-scalar entry on the
    oracle-converted scalars (the byte path is pinned at this size against
    the oracle above)."""
    n = 1 << 20
    pts = hip.synth_points(n, 0x51)
    rnd = np.random.default_rng(3)
    fr = rnd.integers(0, 1 << 63, size=(n, 4), dtype=np.uint64)
    fr[:, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)      # most values below r, a few above
    fr[::4096, 3] |= np.uint64(0xF000000000000000)
    sc = orc.fr_to_scalars(fr)
    ok, got = hip.p1s_mult_fr(pts, fr, "pippenger")
    assert ok
    assert hip.compress(got) == hip.compress(hip.p1s_mult_pippenger(pts, sc, 255))


# ---------------------------------------------------------------- batch normalisation

def _jacobian_inputs(orc, n, seed, inf_every=0):
    """n Jacobian points (random multiples of G at random Z), with Z = 0
    every inf_every-th point."""
    rnd = np.random.default_rng(seed)
    aff = orc.gen_points(n, seed, threads=8)
    if inf_every:
        aff[::inf_every] = 0
    lam = np.zeros((n, 6), np.uint64)
    lam[:, :5] = rnd.integers(1, 1 << 63, size=(n, 5), dtype=np.uint64)   # nonzero, below p
    return aff, orc.p1s_from_affine_scaled(aff, lam)


@pytest.mark.parametrize("n", [1, 2, 3, 100, 1536, 1537, 65536, 65537, 200001, (1 << 19) + 12345])
def test_to_affine_matches_oracle(hip, orc, n):
    """From 2^19 points the host-buffer entry moves the points in ranges with
    uploads and downloads overlapping; every size against the oracle."""
    aff, jac = _jacobian_inputs(orc, n, 0xAF + n, inf_every=13)
    got = hip.p1s_to_affine(jac)
    assert np.array_equal(got, orc.p1s_to_affine(jac))
    assert np.array_equal(got, aff)


@pytest.mark.parametrize("mode", ["exact", "mixed"])
def test_to_affine_descriptors(hip, orc, mode):
    aff, jac = _jacobian_inputs(orc, 300, 0xD5, inf_every=7)
    assert np.array_equal(hip.p1s_to_affine(jac, mode=mode, k=37), aff)


def test_to_affine_all_infinity_and_mixed_z(hip, orc):
    jac = np.zeros((70000, 18), np.uint64)
    assert not hip.p1s_to_affine(jac).any()
    g = orc.from_affine(orc.generator_affine())
    jac[5] = orc.p1_mult(g, (12345).to_bytes(32, "little"), 255)   # Z from the formulas
    jac[69999] = g                                                  # Z = one
    got = hip.p1s_to_affine(jac)
    assert np.array_equal(got[5], orc.to_affine(jac[5])) and np.array_equal(got[69999], orc.generator_affine())
    assert not np.delete(got, [5, 69999], axis=0).any()


def test_to_affine_inversion_edge_values(hip, orc):
    """One point per lane, so each lane inverts its point's Z word image
    itself: Z images 1, 2, p - 1, p - 2, (p +- 1)/2, 2^k, p - 2^k and 2^k - 1
    (the values a binary gcd inversion handles last) give back the affine
    points."""
    P = pyref.P
    zs = [1, 2, 3, P - 1, P - 2, (P - 1) // 2, (P + 1) // 2]
    zs += [1 << k for k in range(0, 381, 7)] + [P - (1 << k) for k in range(0, 381, 7)]
    zs += [(1 << k) - 1 for k in range(2, 381, 11)]
    n = len(zs)
    aff = orc.gen_points(n, 0x2E, threads=8)
    lam = np.array([[(z >> (64 * j)) & (2**64 - 1) for j in range(6)] for z in zs], dtype=np.uint64)
    jac = orc.p1s_from_affine_scaled(aff, lam)
    got = hip.p1s_to_affine(jac)
    assert np.array_equal(got, orc.p1s_to_affine(jac))
    assert np.array_equal(got, aff)


def test_to_affine_device_2pow20_round_trip(hip, orc):
    """At 2^20 points: affine -> scaled Jacobian -> the device kernel gives
    back the same affine points (a size-independent round trip)."""
    n = 1 << 20
    aff = hip.synth_points(n, 0x77)
    aff[::4099] = 0
    rnd = np.random.default_rng(5)
    lam = np.zeros((n, 6), np.uint64)
    lam[:, :5] = rnd.integers(1, 1 << 63, size=(n, 5), dtype=np.uint64)
    jac = orc.p1s_from_affine_scaled(aff, lam)
    d_in = hip.DeviceBuffer(jac)
    d_out = hip.DeviceBuffer(np.zeros((n, 12), np.uint64))
    try:
        hip.p1s_to_affine_device(d_out.ptr, d_in.ptr, n)
        got = d_out.read().view(np.uint64).reshape(n, 12)
    finally:
        d_in.close()
        d_out.close()
    assert np.array_equal(got, aff)
    assert np.array_equal(hip.p1s_to_affine(jac), aff)


# ---------------------------------------------------------------- fixed-base tables

@pytest.mark.parametrize("wbits", [2, 3, 4, 8, 10, 14])
def test_wbits_table_matches_oracle(hip, orc, wbits):
    n = 3 if wbits == 14 else 40
    pts = orc.gen_points(n, 0x7B + wbits, threads=8)
    pts[1] = 0
    want = orc.wbits_precompute(pts, wbits)
    assert np.array_equal(hip.wbits_precompute(pts, wbits), want)


@pytest.mark.parametrize("wbits", [2, 3, 5, 8, 11, 14])
@pytest.mark.parametrize("nbits", [1, 8, 64, 255, 256])
def test_wbits_msm_matches_oracle(hip, orc, wbits, nbits):
    n = 3 if wbits == 14 else 37
    rnd = random.Random(wbits * 1000 + nbits)
    pts = orc.gen_points(n, 0x5A + wbits, threads=8)
    pts[2] = 0
    nb = (nbits + 7) // 8
    S = _scal([rnd.getrandbits(8 * nb) for _ in range(n)], nb)   # bits above nbits are ignored
    S[0] = 0
    tab = orc.wbits_precompute(pts, wbits)
    want = orc.compress(orc.mult_wbits(tab, wbits, S, n, nbits))
    assert want == orc.compress(orc.msm_pippenger(pts, S, n, nbits))
    assert hip.compress(hip.mult_wbits(tab, wbits, S, nbits)) == want


def test_wbits_known_answers(hip, orc):
    """C_api_tests.cpp:1606-1690: nbits 8, wbits 4 and 10."""
    g = orc.from_affine(orc.generator_affine())
    for ks, sc, wb, total in (([3, 5, 7], [2, 3, 5], 4, 56), ([3, 5, 7, 9, 11], [2, 3, 5, 7, 11], 10, 240)):
        pts = np.stack([orc.to_affine(orc.p1_mult(g, k.to_bytes(32, "little"), 255)) for k in ks])
        tab = hip.wbits_precompute(pts, wb)
        r = hip.mult_wbits(tab, wb, np.array(sc, np.uint8).reshape(-1, 1), 8)
        assert hip.compress(r) == pyref.compress(pyref.mul(total, pyref.G))


@pytest.mark.parametrize("mode", ["exact", "mixed"])
def test_wbits_descriptors_and_packed_tail(hip, orc, mode):
    n, wbits = 21, 5
    pts = orc.gen_points(n, 0x33, threads=8)
    sc = orc.gen_scalars(n, 0x34)
    want_tab = orc.wbits_precompute(pts, wbits)
    assert np.array_equal(hip.wbits_precompute(pts, wbits, mode=mode, k=6), want_tab)
    want = orc.compress(orc.msm_pippenger(pts, sc, n, 255))
    assert hip.compress(hip.mult_wbits(want_tab, wbits, sc, 255, mode=mode, k=6)) == want
    # the points packed into the table's last n entries (an allowed overlap)
    L = hip.load()
    m = 1 << (wbits - 1)
    table = np.zeros((n * m, 12), np.uint64)
    table[n * m - n:] = pts
    heads = np.array([table[n * m - n:].ctypes.data, 0], np.uint64)
    L.blst_p1s_mult_wbits_precompute(hip._ptr(table), wbits, hip._ptr(heads), n)
    assert hip.last_error() == ""
    assert np.array_equal(table, want_tab)


def test_wbits_device_resident_4096(hip, orc):
    """A resident 4096-point setup: the device table's rows equal the
    oracle's (sampled) and the device MSM over it equals the plain MSM."""
    n, wbits = 4096, 10
    m = 1 << (wbits - 1)
    pts = hip.generator_multiples(hip.synth_srs_scalars(n, 0x44))
    sc = hip.synth_scalars(n, 0x45)
    d_pts = hip.DeviceBuffer(pts)
    d_tab = hip.DeviceAlloc(n * m * 96)
    d_sc = hip.DeviceBuffer(sc)
    try:
        hip.wbits_precompute_device(d_tab.ptr, wbits, d_pts.ptr, n)
        got = hip.mult_wbits_device(d_tab.ptr, wbits, n, d_sc.ptr, 255)
        plain = hip.msm_device(d_pts.ptr, d_sc.ptr, n, 255)
        tab = d_tab.read().view(np.uint64).reshape(n * m, 12)
    finally:
        d_pts.close()
        d_tab.close()
        d_sc.close()
    assert hip.compress(got) == hip.compress(plain)
    for i in (0, 1, 2047, 4095):
        assert np.array_equal(tab[i * m:(i + 1) * m], orc.wbits_precompute(pts[i:i + 1], wbits))
    assert hip.compress(hip.mult_wbits(tab, wbits, sc, 255)) == hip.compress(plain)


def test_wbits_device_2pow16_property(hip):
    """At 2^16 points with an 8-bit table the device MSM equals the plain MSM
    (a size-independent check against the path pinned above)."""
    n, wbits = 1 << 16, 8
    pts = hip.synth_points(n, 0x61)
    sc = hip.synth_scalars(n, 0x62)
    d_pts = hip.DeviceBuffer(pts)
    d_tab = hip.DeviceAlloc(n * (1 << (wbits - 1)) * 96)
    d_sc = hip.DeviceBuffer(sc)
    try:
        hip.wbits_precompute_device(d_tab.ptr, wbits, d_pts.ptr, n)
        got = hip.mult_wbits_device(d_tab.ptr, wbits, n, d_sc.ptr, 255)
        plain = hip.msm_device(d_pts.ptr, d_sc.ptr, n, 255)
    finally:
        d_pts.close()
        d_tab.close()
        d_sc.close()
    assert hip.compress(got) == hip.compress(plain)


# ---------------------------------------------------------------- G2

import pyref2  # noqa: E402


@pytest.fixture(scope="module")
def g2_points(orc):
    return orc.gen_p2_points(2048, 0x6262)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 33, 100, 513, 2048])
def test_g2_msm_random(hip, orc, g2_points, n):
    pts = np.ascontiguousarray(g2_points[:n])
    sc = orc.gen_scalars(n, 2000 + n)
    want = orc.p2_compress(orc.msm_p2_pippenger(pts, sc, n, 255))
    assert orc.p2_compress(hip.p2s_mult_pippenger(pts, sc, 255)) == want


@pytest.mark.parametrize("nbits", [1, 8, 64, 255, 256])
@pytest.mark.parametrize("window", [0, 1, 4, 9, 13])
def test_g2_msm_nbits_windows(hip, orc, g2_points, nbits, window):
    n = 150
    rnd = random.Random(nbits * 100 + window)
    nb = (nbits + 7) // 8
    S = _scal([rnd.getrandbits(8 * nb) for _ in range(n)], nb)
    pts = np.ascontiguousarray(g2_points[:n])
    want = orc.p2_compress(orc.msm_p2_pippenger(pts, S, n, nbits, window))
    assert orc.p2_compress(hip.p2s_mult_pippenger(pts, S, nbits, window=window)) == want


def test_g2_edge_inputs(hip, orc, g2_points):
    """Zero scalars, infinity points, P and -P with equal scalars, repeated
    points (the doubling path) and r - 1."""
    n = 64
    pts = np.ascontiguousarray(g2_points[:n]).copy()
    rnd = random.Random(5)
    ks = [rnd.getrandbits(255) for _ in range(n)]
    ks[0] = 0
    pts[1] = 0
    neg = orc.p2_to_affine(orc.p2_mult(orc.p2_from_affine(pts[2]), (R - 1).to_bytes(32, "little"), 255))
    pts[3] = neg
    ks[3] = ks[2]
    for i in range(10, 20):
        pts[i] = pts[9]
        ks[i] = ks[9]
    ks[30] = R - 1
    ks[31] = 1
    S = _scal(ks)
    want = orc.p2_compress(orc.msm_p2_pippenger(pts, S, n, 255))
    assert orc.p2_compress(hip.p2s_mult_pippenger(pts, S, 255)) == want
    # everything cancels or is zero: infinity
    Z = _scal([0] * n)
    assert orc.p2_compress(hip.p2s_mult_pippenger(pts, Z, 255)) == bytes([0xC0]) + bytes(95)


@pytest.mark.parametrize("n", [(1 << 19) + 333, (1 << 20) + 64])
def test_g2_host_buffers_in_point_ranges(hip, orc, n):
    """From 2^19 points the G2 host-buffer entry copies the points in ranges
    that share one set of buckets; it equals the one-range MSM over the same
    inputs in device memory (zero scalars and infinity points included)."""
    pts = hip.synth_p2_points(n, 0x3C + n)
    sc = hip.synth_scalars(n, 0x3D + n)
    for i in (0, n // 2, n - 1):
        sc[i] = 0
        pts[i + 3 if i + 3 < n else i - 3] = 0
    dp, ds = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    try:
        want = hip.p2_msm_device(dp.ptr, ds.ptr, n, 255)
    finally:
        dp.close()
        ds.close()
    assert orc.p2_compress(hip.p2s_mult_pippenger(pts, sc, 255)) == orc.p2_compress(want)


@pytest.mark.parametrize("entry", FR_ENTRIES)
def test_g2_fr_known_answer_ten_g(hip, orc, entry):
    """Scalars 1..4 in Montgomery form on four copies of the G2 generator
    give [10]G2; a scratch one byte short is refused with ret untouched
    (BindingHeader_tests.cpp:642-690 on G2)."""
    g = orc.p2_generator_affine()
    pts = np.ascontiguousarray(np.tile(g, (4, 1)))
    fr = orc.fr_from_values([1, 2, 3, 4])
    ok, ret = hip.p2s_mult_fr(pts, fr, entry)
    assert ok
    want = orc.p2_mult(orc.p2_from_affine(g), (10).to_bytes(1, "little"), 8)
    assert orc.p2_compress(ret) == orc.p2_compress(want)
    short = np.full(36, 7, np.uint64)
    ok, out = hip.p2s_mult_fr(pts, fr, entry, scratch_size=hip.p2_fr_scratch_sizeof(4, entry) - 1, ret=short)
    assert not ok and (out == 7).all()


@pytest.mark.parametrize("entry", FR_ENTRIES)
@pytest.mark.parametrize("n", [1, 3, 100, 2048])
def test_g2_fr_random_matches_oracle(hip, orc, g2_points, entry, n):
    """Random Montgomery-form scalars with 0, one, r - 1, r, 2^256 - 1 among
    them, against the oracle's conversion and G2 MSM."""
    rnd = random.Random(n * 11 + len(entry))
    pts = np.ascontiguousarray(g2_points[:n])
    vals = [rnd.randrange(R) for _ in range(n)]
    for k, v in enumerate([0, 1, R - 1]):
        if k < n:
            vals[k] = v
    fr = orc.fr_from_values(vals)
    if n > 8:
        fr[5] = np.uint64(0xFFFFFFFFFFFFFFFF)
        fr[6] = orc.limbs(R, 4)
    want = orc.p2_compress(orc.msm_p2_pippenger(pts, orc.fr_to_scalars(fr), n, 255))
    ok, ret = hip.p2s_mult_fr(pts, fr, entry)
    assert ok and orc.p2_compress(ret) == want


def test_g2_fr_capacity_scratch_2pow16(hip, orc):
    """At 2^16 points (past the reference's G2 batch-affine range) the
    blst_p2s_mult_fr entry with a capacity-sized scratch for 2^17 equals the
    byte-scalar G2 entry on the oracle-converted scalars."""
    n = 1 << 16
    pts = hip.synth_p2_points(n, 0x77)
    rng = np.random.default_rng(8)
    fr = rng.integers(0, 1 << 63, size=(n, 4), dtype=np.uint64)
    fr[:, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)
    fr[::1024, 3] |= np.uint64(0xF000000000000000)
    sc = orc.fr_to_scalars(fr)
    ok, got = hip.p2s_mult_fr(pts, fr, "msm", scratch_size=hip.p2_fr_scratch_sizeof(1 << 17, "msm"))
    assert ok
    assert orc.p2_compress(got) == orc.p2_compress(hip.p2s_mult_pippenger(pts, sc, 255))


@pytest.mark.parametrize("wbits", [2, 4, 8])
def test_g2_wbits_table_matches_oracle(hip, orc, g2_points, wbits):
    """blst_p2s_mult_wbits_precompute writes the oracle's table bit for bit
    (canonical affine rows of [1]P .. [2^(wbits-1)]P; an infinity point gives
    an all-zero row)."""
    pts = np.ascontiguousarray(g2_points[:37]).copy()
    pts[5] = 0
    got = hip.p2_wbits_precompute(pts, wbits)
    assert np.array_equal(got, orc.p2_wbits_precompute(pts, wbits))


@pytest.mark.parametrize("wbits,nbits", [(2, 255), (4, 64), (5, 256), (8, 255), (13, 17)])
def test_g2_wbits_msm_matches_oracle(hip, orc, g2_points, wbits, nbits):
    """blst_p2s_mult_wbits over the oracle's table equals the oracle's table
    MSM and its G2 Pippenger MSM, with zero scalars, infinity and P, -P."""
    n = 300
    pts = np.ascontiguousarray(g2_points[:n]).copy()
    pts[3] = 0
    pts[7] = orc.p2_to_affine(orc.p2_mult(orc.p2_from_affine(pts[6]), (R - 1).to_bytes(32, "little"), 255))
    rnd = random.Random(wbits * 1000 + nbits)
    nb = (nbits + 7) // 8
    ks = [rnd.getrandbits(nbits) for _ in range(n)]
    ks[0] = 0
    ks[7] = ks[6]
    S = _scal(ks, nb)
    tab = orc.p2_wbits_precompute(pts, wbits)
    want = orc.p2_compress(orc.msm_p2_pippenger(pts, S, n, nbits))
    assert orc.p2_compress(orc.p2_mult_wbits(tab, wbits, S, n, nbits)) == want
    assert orc.p2_compress(hip.p2_mult_wbits(tab, wbits, S, nbits)) == want


@pytest.mark.parametrize("mode", ["exact", "mixed"])
def test_g2_wbits_descriptors(hip, orc, g2_points, mode):
    n = 40
    pts = np.ascontiguousarray(g2_points[:n])
    S = orc.gen_scalars(n, 0x3A)
    tab = hip.p2_wbits_precompute(pts, 4, mode=mode, k=7)
    assert np.array_equal(tab, orc.p2_wbits_precompute(pts, 4))
    want = orc.p2_compress(orc.msm_p2_pippenger(pts, S, n, 255))
    assert orc.p2_compress(hip.p2_mult_wbits(tab, 4, S, 255, mode=mode, k=7)) == want


@pytest.mark.parametrize("n", [1, 5, 1000, 1 << 16])
def test_g2_to_affine_matches_oracle(hip, orc, g2_points, n):
    """blst_p2s_to_affine against the oracle's batch normalisation, with
    Z = 0 points (all-zero output) among Jacobian inputs of varied Z."""
    G = orc.p2_from_affine(orc.p2_generator_affine())
    base = np.stack([orc.p2_mult(G, (1000003 * i + 7).to_bytes(8, "little"), 64) for i in range(64)])
    jac = np.ascontiguousarray(np.tile(base, (-(-n // 64), 1))[:n]).copy()
    jac[::17] = 0
    got = hip.p2s_to_affine(jac)
    assert np.array_equal(got, orc.p2s_to_affine(jac))
    if n == 1000:
        assert np.array_equal(hip.p2s_to_affine(jac, mode="exact"), got)


def test_g2_hash_to_curve_points_textbook(hip):
    """The 10 G2 hash-to-curve outputs as bases with small scalars, against
    the textbook Fp2 model."""
    vs = json.load(open(os.path.join(GOLD, "h2c_g2.json")))
    pts = [(tuple(int(c, 16) for c in v["x"]), tuple(int(c, 16) for c in v["y"])) for v in vs]
    ks = [3 + 7 * i for i in range(len(pts))]
    limbs = np.array([pyref2.affine_to_limbs(p) for p in pts], np.uint64)
    got = hip.p2s_mult_pippenger(limbs, _scal(ks, 1), 8)
    assert pyref2.jacobian_limbs_to_affine(list(got)) == pyref2.msm(pts, ks)


@pytest.mark.parametrize("mode", ["exact", "mixed"])
def test_g2_descriptors_and_tile(hip, orc, g2_points, mode):
    n = 45
    pts = np.ascontiguousarray(g2_points[:n])
    sc = orc.gen_scalars(n, 77)
    want = orc.p2_compress(orc.msm_p2_pippenger(pts, sc, n, 255))
    assert orc.p2_compress(hip.p2s_mult_pippenger(pts, sc, 255, mode=mode, k=11)) == want
    for bit0, window in ((0, 8), (248, 8), (100, 13)):
        t = hip.p2s_mult_pippenger(pts, sc, 255, window=window, bit0=bit0)
        assert orc.p2_compress(t) == orc.p2_compress(orc.msm_p2_tile(pts, sc, n, 255, bit0, window))


def test_g2_device_linearity_2pow15(hip, orc):
    """MSM(k) + MSM(k') == MSM(k + k' mod r) over 2^15 resident G2 bases."""
    n = 1 << 15
    pts = orc.gen_p2_points(n, 0x7777)
    a = hip.synth_scalars(n, 0x71)
    b = hip.synth_scalars(n, 0x72)
    s = np.zeros_like(a)
    for i in range(n):
        v = (int.from_bytes(a[i].tobytes(), "little") + int.from_bytes(b[i].tobytes(), "little")) % R
        s[i] = np.frombuffer(v.to_bytes(32, "little"), np.uint8)
    dp = hip.DeviceBuffer(pts)
    da, db, ds = hip.DeviceBuffer(a), hip.DeviceBuffer(b), hip.DeviceBuffer(s)
    try:
        ra = hip.p2_msm_device(dp.ptr, da.ptr, n, 255)
        rb = hip.p2_msm_device(dp.ptr, db.ptr, n, 255)
        rs = hip.p2_msm_device(dp.ptr, ds.ptr, n, 255)
    finally:
        for d in (dp, da, db, ds):
            d.close()
    assert orc.p2_compress(orc.p2_add(ra, rb)) == orc.p2_compress(rs)
    assert orc.p2_compress(rs) != bytes([0xC0]) + bytes(95)


# ---------------------------------------------------------------- comb tables

@pytest.mark.parametrize("wbits,nbits,n,njobs", [(8, 255, 512, 5), (3, 64, 100, 3), (11, 256, 64, 2), (2, 1, 9, 2)])
def test_comb_batch_matches_plain(hip, orc, wbits, nbits, n, njobs):
    """Batched MSMs over a comb table equal the plain MSM of each job (and the
    oracle for the first)."""
    pts = hip.generator_multiples(hip.synth_srs_scalars(n, 0x90 + n))
    pts[min(3, n - 1)] = 0
    nb = (nbits + 7) // 8
    rnd = random.Random(wbits * 7 + nbits)
    S = _scal([rnd.getrandbits(8 * nb) for _ in range(n * njobs)], nb).reshape(njobs, n, nb)
    S[0, 0] = 0
    size = hip.comb_table_bytes(n, wbits, nbits)
    d_pts = hip.DeviceBuffer(pts)
    d_tab = hip.DeviceAlloc(size)
    d_sc = hip.DeviceBuffer(S)
    try:
        hip.comb_table_build(d_tab.ptr, wbits, d_pts.ptr, n, nbits)
        got = hip.comb_msm_batch(d_tab.ptr, wbits, n, d_sc.ptr, njobs, nbits)
        plain = [hip.msm_device(d_pts.ptr, d_sc.ptr + k * n * nb, n, nbits) for k in range(njobs)]
    finally:
        d_pts.close()
        d_tab.close()
        d_sc.close()
    for k in range(njobs):
        assert hip.compress(got[k]) == hip.compress(plain[k]), k
    assert hip.compress(got[0]) == orc.compress(orc.msm_pippenger(pts, S[0], n, nbits))
