"""The batch-affine probe (test kit) computes the sums it times: every
sampled pair's batch-affine result equals the oracle's P + Q (affine,
canonical Montgomery limbs), and its XYZZ result names the same point, for
contiguous and gathered pairs and every batch shape.  The probe's timings
are reported by bench.py; here only the values are checked."""
import numpy as np
import pytest

import pyref

pytestmark = pytest.mark.gpu
P = pyref.P
R = 1 << 384


def _val(limbs):
    return sum(int(x) << (64 * i) for i, x in enumerate(limbs))


@pytest.mark.parametrize("K", [8, 32, 128])
@pytest.mark.parametrize("gather", [False, True])
def test_probe_sums_match_oracle(hip, orc, K, gather):
    npairs = 1 << 15
    pts = hip.synth_points(2 * npairs, 0xBA00 + K)
    d_p = hip.DeviceBuffer(pts)
    d_i = None
    rng = np.random.default_rng(K + 7 * gather)
    idx = None
    if gather:
        idx = rng.permutation(2 * npairs).astype(np.uint32)
        d_i = hip.DeviceBuffer(idx)
    try:
        print("probe K=%d gather=%s" % (K, gather), flush=True)
        ns, aff, xz = hip.batch_affine_probe(d_p.ptr, npairs, K, d_i.ptr if d_i else None)
    finally:
        d_p.close()
        if d_i:
            d_i.close()
    assert ns[0] > 0 and ns[1] > 0
    sample = sorted(set(rng.integers(0, npairs, 600).tolist()) | {0, npairs - 1})
    for j in sample:
        a, b = (int(idx[2 * j]), int(idx[2 * j + 1])) if gather else (2 * j, 2 * j + 1)
        want = orc.to_affine(orc.p1_add(orc.from_affine(pts[a]), orc.from_affine(pts[b])))
        assert np.array_equal(aff[j], want), (K, gather, j)
        # the XYZZ sum names the same point: X = x ZZ, Y = y ZZZ (Montgomery
        # words: x_m ZZ_m = X_m R mod p)
        X, Y, ZZ, ZZZ = (_val(xz[j, 6 * c:6 * c + 6]) for c in range(4))
        x, y = _val(want[:6]), _val(want[6:])
        assert x * ZZ % P == X * R % P and y * ZZZ % P == Y * R % P, (K, gather, j)


@pytest.mark.parametrize("stride", [96, 128])
def test_gather_probe_sums_match_oracle(hip, orc, stride):
    """The gather probe's bucket sums (32 randomly gathered points per lane,
    with the accumulation kernel's arithmetic) equal the oracle's sums, read
    from the 96-byte array and from the 128-byte padded copy, with repeated
    points, a point and its negation's absence, and infinity among them."""
    n, entries = 1 << 12, 1 << 15
    pts = hip.synth_points(n, 0x6A7)
    pts[5] = 0                                    # infinity
    rng = np.random.default_rng(stride)
    idx = rng.integers(0, n, entries).astype(np.uint32)
    lanes = entries // 32
    idx[:64] = 7                                  # lanes 0..63: the same point twice (a doubling)
    idx[lanes:lanes + 64] = 7
    src = pts
    if stride == 128:
        src = np.zeros((n, 16), np.uint64)
        src[:, :12] = pts
    d_p, d_i = hip.DeviceBuffer(src), hip.DeviceBuffer(idx)
    try:
        print("gather probe stride %d" % stride, flush=True)
        ns, sums = hip.gather_probe(d_p.ptr, stride, d_i.ptr, entries)
    finally:
        d_p.close()
        d_i.close()
    assert ns > 0
    for g in sorted(set(rng.integers(0, lanes, 200).tolist()) | {0, 1, lanes - 1}):
        sel = idx[np.arange(32) * lanes + g]
        want = orc.msm_pippenger(pts[sel], np.ones((32, 1), np.uint8), 32, 1)
        X, Y, ZZ, ZZZ = (_val(sums[g, 6 * c:6 * c + 6]) for c in range(4))
        if ZZ == 0:
            assert orc.compress(want) == orc.compress(np.zeros(18, np.uint64))
            continue
        aff = orc.to_affine(want)
        x, y = _val(aff[:6]), _val(aff[6:])
        assert x * ZZ % P == X * R % P and y * ZZZ % P == Y * R % P, (stride, g)
