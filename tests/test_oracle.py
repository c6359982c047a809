"""Pins the oracle (oracle/blst_oracle.c) against checks that do not use it:
Python big-integer field arithmetic, the textbook affine group law
(tests/pyref.py), the IETF generator encoding, the hash-to-curve G1 vectors
and the reference's MSM fuzz corpus (tests/golden/).  CPU only."""
import json
import os
import random

import numpy as np
import pytest

import pyref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = pyref.P


def mont(x):
    return np.array(pyref.fp_limbs(pyref.to_mont(x)), dtype=np.uint64)


def unmont(a):
    return pyref.from_mont(pyref.limbs_value(a))


def aff_limbs(pt):
    return np.array(pyref.affine_to_limbs(pt), dtype=np.uint64)


EDGE = [0, 1, 2, P - 1, P - 2, (P - 1) // 2, (P + 1) // 2, (1 << 381) % P, (1 << 383) % P]


def test_fp_ops_against_bigint(orc):
    rnd = random.Random(11)
    vals = EDGE + [rnd.randrange(P) for _ in range(200)]
    for i, a in enumerate(vals):
        b = vals[(7 * i + 3) % len(vals)]
        am, bm = mont(a), mont(b)
        assert unmont(orc.fp_mul(am, bm)) == a * b % P
        assert unmont(orc.fp_binop("or_fp_add", am, bm)) == (a + b) % P
        assert unmont(orc.fp_binop("or_fp_sub", am, bm)) == (a - b) % P
        assert unmont(orc.fp_unop("or_fp_neg", am)) == (-a) % P
        # raw Montgomery product of the limb images: a*b*2^-384
        raw = orc.fp_mul(orc.limbs(a), orc.limbs(b))
        assert orc.value(raw) == a * b * pow(1 << 384, -1, P) % P
        if a:
            assert unmont(orc.fp_unop("or_fp_inv", am)) == pow(a, -1, P)
    assert unmont(orc.fp_unop("or_fp_inv", mont(0))) == 0


def test_montgomery_conversions(orc):
    rnd = random.Random(12)
    for a in EDGE + [rnd.randrange(P) for _ in range(50)]:
        m = orc.fp_unop("or_fp_to_mont", orc.limbs(a))
        assert orc.value(m) == a * (1 << 384) % P
        assert orc.value(orc.fp_unop("or_fp_from_mont", m)) == a


def test_generator_and_infinity_encoding(orc):
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    g = orc.generator_affine()
    assert orc.on_curve(g)
    assert pyref.limbs_to_affine(list(g)) == pyref.G
    assert orc.compress(orc.from_affine(g)).hex() == kat["g1_generator_compressed"]
    assert pyref.compress(pyref.G).hex() == kat["g1_generator_compressed"]
    assert orc.compress(np.zeros(18, np.uint64)).hex() == kat["g1_infinity_compressed"]


def test_hash_to_curve_points(orc):
    pts = json.load(open(os.path.join(GOLD, "h2c_g1.json")))
    assert len(pts) == 10
    for v in pts:
        pt = (int(v["x"], 16), int(v["y"], 16))
        a = aff_limbs(pt)
        assert orc.on_curve(a)
        assert orc.compress(orc.from_affine(a)).hex() == v["compressed"]
        # a Jacobian representative with Z != 1 compresses the same way
        j = orc.from_affine(a)
        j2 = orc.p1_add(j, np.zeros(18, np.uint64))
        assert orc.compress(orc.p1_double(j2)) == pyref.compress(pyref.add(pt, pt))
        assert pyref.decompress(bytes.fromhex(v["compressed"])) == pt


def test_group_law_against_textbook(orc):
    rnd = random.Random(13)
    g = orc.from_affine(orc.generator_affine())
    for _ in range(12):
        a, b = rnd.randrange(1, 1 << 64), rnd.randrange(1, 1 << 64)
        pa = orc.p1_mult(g, a.to_bytes(8, "little"), 64)
        pb = orc.p1_mult(g, b.to_bytes(8, "little"), 64)
        assert orc.compress(pa) == pyref.compress(pyref.mul(a, pyref.G))
        assert orc.compress(orc.p1_add(pa, pb)) == pyref.compress(pyref.mul(a + b, pyref.G))
        assert orc.compress(orc.p1_double(pa)) == pyref.compress(pyref.mul(2 * a, pyref.G))
        assert orc.compress(orc.p1_add(pa, pa)) == pyref.compress(pyref.mul(2 * a, pyref.G))
    # P + (-P) = infinity
    pa = orc.p1_mult(g, (12345).to_bytes(2, "little"), 16)
    neg = pa.copy()
    neg[6:12] = orc.fp_unop("or_fp_neg", pa[6:12])
    assert orc.compress(orc.p1_add(pa, neg)) == pyref.compress(None)


def _pts_and_scalars(n, nbits, seed):
    rnd = random.Random(seed)
    ks = [rnd.randrange(1, pyref.R_ORDER) for _ in range(n)]
    pts = [pyref.mul(k, pyref.G) for k in ks]
    nb = (nbits + 7) // 8
    ss = [rnd.getrandbits(nbits) for _ in range(n)]
    S = np.array([list(s.to_bytes(nb, "little")) for s in ss], dtype=np.uint8).reshape(n, nb)
    A = np.array([pyref.affine_to_limbs(p) for p in pts], dtype=np.uint64)
    return pts, ss, A, S


@pytest.mark.parametrize("n,nbits", [(2, 8), (3, 8), (5, 1), (7, 17), (9, 64), (12, 255), (6, 256)])
def test_msm_against_textbook(orc, n, nbits):
    pts, ss, A, S = _pts_and_scalars(n, nbits, 100 * n + nbits)
    want = pyref.compress(pyref.msm(pts, ss))
    assert orc.compress(orc.msm_naive(A, S, n, nbits)) == want
    assert orc.compress(orc.msm_pippenger(A, S, n, nbits)) == want
    for w in (1, 2, 3, 5, 8):
        assert orc.compress(orc.msm_pippenger(A, S, n, nbits, window=w)) == want
    assert orc.compress(orc.msm_pippenger(A, S, n, nbits, threads=3)) == want


def test_msm_batch_against_textbook(orc):
    """The job-parallel MSM that batched GPU launches are checked against: every job of a
    ragged batch (empty, one point, short and long scalars) equals the
    textbook sum, in any job order and thread count."""
    specs = [(12, 255), (0, 255), (1, 256), (5, 1), (9, 64), (7, 17), (11, 255), (3, 8)]
    jobs, want = [], []
    for k, (n, nbits) in enumerate(specs):
        pts, ss, A, S = _pts_and_scalars(max(n, 1), nbits, 7000 + k)
        jobs.append((A, S, n, nbits))
        want.append(pyref.compress(pyref.msm(pts[:n], ss[:n])))
    for threads in (1, 3, 16):
        got = orc.msm_batch(jobs, threads=threads)
        assert [orc.compress(g) for g in got] == want, threads


def test_msm_closed_form_1G_plus_2_2G(orc):
    # 1*G + 2*(2G) = 5G (C_api_tests.cpp:1488-1515 pattern)
    pts = [pyref.G, pyref.mul(2, pyref.G)]
    A = np.array([pyref.affine_to_limbs(p) for p in pts], dtype=np.uint64)
    S = np.array([[1], [2]], dtype=np.uint8)
    assert orc.compress(orc.msm_pippenger(A, S, 2, 8)) == pyref.compress(pyref.mul(5, pyref.G))


def test_scalars_beyond_r_act_mod_r(orc):
    # Group_scalar_mult_tests.cpp:757-808: two copies of G with 2^256-1
    A = np.array([pyref.affine_to_limbs(pyref.G)] * 2, dtype=np.uint64)
    S = np.full((2, 32), 0xFF, np.uint8)
    want = pyref.compress(pyref.mul(2 * ((1 << 256) - 1) % pyref.R_ORDER, pyref.G))
    assert orc.compress(orc.msm_pippenger(A, S, 2, 256)) == want


def test_msm_edge_inputs(orc):
    r = pyref.R_ORDER
    pts = [pyref.mul(k, pyref.G) for k in (3, 5, 7, 11, 13)]
    cases = [
        ([0, 0, 0, 0, 0], None),
        ([1, 0, 0, 0, 0], pts[0]),
        ([r - 1, 0, 0, 0, 0], pyref.neg(pts[0])),
    ]
    for ss, want in cases:
        A = np.array([pyref.affine_to_limbs(p) for p in pts], dtype=np.uint64)
        S = np.array([list(s.to_bytes(32, "little")) for s in ss], dtype=np.uint8)
        assert orc.compress(orc.msm_pippenger(A, S, 5, 255)) == pyref.compress(want)
    # P and -P with equal scalars cancel; repeated points double; infinity skipped
    plist = [pts[0], pyref.neg(pts[0]), pts[1], pts[1], None]
    ss = [77, 77, 5, 5, 123]
    A = np.array([pyref.affine_to_limbs(p) for p in plist], dtype=np.uint64)
    S = np.array([list(s.to_bytes(32, "little")) for s in ss], dtype=np.uint8)
    want = pyref.mul(10, pts[1])
    assert orc.compress(orc.msm_pippenger(A, S, 5, 255)) == pyref.compress(want)


def test_fuzz_corpus_replays(orc):
    cases = json.load(open(os.path.join(GOLD, "fuzz_msm.json")))
    assert len(cases) == 18
    for c in cases:
        n = c["n"]
        A = np.array([pyref.affine_to_limbs(pyref.mul(m, pyref.G)) for m in c["point_multipliers"]],
                     dtype=np.uint64)
        S = np.array(c["scalars"], dtype=np.uint8).reshape(n, 1)
        assert orc.compress(orc.msm_pippenger(A, S, n, 8)).hex() == c["expected"]
        assert orc.compress(orc.msm_pippenger(A, S, n, 8, window=8)).hex() == c["expected"]
        T = np.array(c["tile_scalars"], dtype=np.uint8).reshape(n, 1)
        assert orc.compress(orc.msm_tile(A, T, n, 8, 0, 8)).hex() == c["tile_expected"]


def test_booth_digits_reconstruct_scalar(orc):
    rnd = random.Random(14)
    for nbits in (1, 2, 7, 8, 9, 31, 64, 100, 255, 256):
        nb = (nbits + 7) // 8
        for _ in range(6):
            k = rnd.getrandbits(nbits) | (rnd.getrandbits(1) << (nbits - 1))
            kb = k.to_bytes(nb, "little")
            for c in (1, 2, 3, 4, 5, 8, 13, 16, 17, 21, 24):
                nwin = 1 + nbits // c
                total = 0
                for w in range(nwin):
                    e = min(nbits - w * c + 1, c)
                    raw = orc.booth_digit(kb, nbits, w * c, c)
                    mag = raw & ((1 << e) - 1)
                    sign = (raw >> e) & 1
                    d = -mag if sign else mag
                    assert abs(d) <= 1 << (e - 1)
                    total += d << (c * w)
                assert total == k, (nbits, c, k)


def test_booth_digits_exhaustive_small(orc):
    for nbits in (1, 2, 3, 4, 5, 6):
        for k in range(1 << nbits):
            for c in (1, 2, 3, 4):
                total = 0
                for w in range(1 + nbits // c):
                    e = min(nbits - w * c + 1, c)
                    raw = orc.booth_digit(bytes([k]), nbits, w * c, c)
                    mag, sign = raw & ((1 << e) - 1), (raw >> e) & 1
                    total += (-mag if sign else mag) << (c * w)
                assert total == k


def test_window_and_scratch_sizes(orc):
    # reference values (multi_scalar.h:329-348, 460-479)
    assert orc.scratch_sizeof(0) == 0
    assert orc.scratch_sizeof(1 << 6) == 3072
    assert orc.scratch_sizeof(4096) == 98304
    assert orc.scratch_sizeof(1 << 16) == 786432
    assert orc.scratch_sizeof(1 << 18) == 3145728
    assert orc.scratch_sizeof(1 << 20) == 12582912
    assert orc.window_size(4096) == 10 and orc.window_size(1 << 20) == 17
    assert orc.window_scratch_sizeof(0) == 0 and orc.window_scratch_sizeof(25) == 0
    assert orc.window_scratch_sizeof(24) == 192 << 23
    for n in (1, 2, 3, 4, 15, 16, 17, 255, 256, 511, 512, 8191, 8192):
        lg = n.bit_length() - 1
        w = lg - 3 if lg > 12 else lg - 2 if lg > 8 else lg - 1 if lg > 4 else (2 if lg else 1)
        assert orc.window_size(n) == w


def test_generated_points_are_multiples_of_g(orc):
    keys = orc.gen_scalars(6, 42)
    pts = orc.gen_points(6, 42, threads=2)
    for k, p in zip(keys, pts):
        assert pyref.limbs_to_affine(list(p)) == pyref.mul(int.from_bytes(k.tobytes(), "little"), pyref.G)
    srs = orc.gen_srs(4, 9, threads=2)
    g = pyref.limbs_to_affine(list(srs[1]))
    assert pyref.limbs_to_affine(list(srs[0])) == pyref.G
    # [tau^2]G from tau: check via discrete structure e.g. srs[2] = tau*srs[1] needs tau;
    # instead check srs[i] are on the curve and distinct
    assert all(pyref.on_curve(pyref.limbs_to_affine(list(s))) for s in srs) and g != pyref.G


def test_edge_case_generator(orc):
    n = 1000
    pts = orc.gen_points(n, 1, threads=2)
    sc = orc.gen_scalars(n, 2)
    p0, s0 = pts.copy(), sc.copy()
    orc.apply_edge_cases(pts, sc, 3)
    r = pyref.R_ORDER
    vals = [int.from_bytes(s.tobytes(), "little") for s in sc]
    assert sum(v == 0 for v in vals) >= 10
    assert sum(v == 1 for v in vals) >= 10
    assert sum(v == r - 1 for v in vals) >= 10
    changed = sum(not np.array_equal(a, b) for a, b in zip(pts, p0))
    assert 1 <= changed <= 5
    assert sum(not np.array_equal(a, b) for a, b in zip(sc, s0)) >= 25


def test_fr_to_scalar_against_bigint(orc):
    """Fr::to_scaler is x * 2^-256 mod r for the Montgomery form x; any
    256-bit x reduces to its canonical class (blst_fr inputs above r)."""
    r = pyref.R_ORDER
    rinv = pow(2, -256, r)
    rnd = random.Random(31)
    xs = [0, 1, 2, r - 1, r, r + 1, (1 << 256) - 1, (1 << 256) % r, (1 << 255)]
    xs += [rnd.getrandbits(256) for _ in range(200)] + [rnd.randrange(r) for _ in range(200)]
    fr = np.stack([orc.limbs(x, 4) for x in xs])
    got = orc.fr_to_scalars(fr)
    for x, g in zip(xs, got):
        assert int.from_bytes(g.tobytes(), "little") == x * rinv % r, hex(x)
    # values -> Montgomery form -> values (blst_fr_from_uint64 then to_scaler)
    vals = [0, 1, 10, r - 1] + [rnd.randrange(r) for _ in range(50)]
    back = orc.fr_to_scalars(orc.fr_from_values(vals))
    assert [int.from_bytes(b.tobytes(), "little") for b in back] == vals


def test_fr_scratch_sizes(orc):
    """c_api.cpp:272-299, 593-627, 758-807, pinned by the reference's own
    binding tests (BindingHeader_tests.cpp:652-653, 1190-1203), by the
    Pippenger bucket size already pinned above and by the batch-affine
    element sizes that C's layout rules give the structs of
    multi_scalar_batch_affine.cpp:16-36 (computed here by ctypes)."""
    import ctypes
    assert orc.pippenger_fr_scratch_sizeof(0) == 0
    assert orc.pippenger_fr_scratch_sizeof(1) == 32
    assert orc.msm_fr_scratch_sizeof(0) == 0
    assert orc.msm_fr_scratch_sizeof(1) == 32
    # Pippenger: 32-byte canonical scalars, then the plain entry's buckets
    for n in (2, 3, 5, 64, 100, 4095, 4096, 4097, 1 << 16, 1 << 20):
        assert orc.pippenger_fr_scratch_sizeof(n) == 32 * n + orc.scratch_sizeof(n), n

    class AffBucket(ctypes.Structure):
        _fields_ = [("x", ctypes.c_uint64 * 6), ("y", ctypes.c_uint64 * 6),
                    ("generation", ctypes.c_uint32), ("queued_epoch", ctypes.c_uint32)]

    class ProjBucket(ctypes.Structure):
        _fields_ = [("point", ctypes.c_uint64 * 18), ("generation", ctypes.c_uint32)]

    class Sched(ctypes.Structure):
        _fields_ = [("point_index", ctypes.c_uint32), ("bucket_index", ctypes.c_uint32),
                    ("subtract", ctypes.c_uint32)]

    def align(v, a):
        return -(-v // a) * a

    def ba(window, batch):
        B = 1 << (window - 1)
        off = B * ctypes.sizeof(AffBucket)
        off = align(off, ctypes.alignment(ProjBucket)) + B * ctypes.sizeof(ProjBucket)
        off = align(off, ctypes.alignment(Sched)) + batch * ctypes.sizeof(Sched)
        off = align(off, 8) + batch * 48
        off = align(off, 8) + batch * 48
        return align(off, 8)

    for n, w, b in ((4096, 10, 128), (2, 1, 128), (5000, 11, 384), (1 << 18, 15, 384), (7, 24, 128)):
        assert orc.batch_affine_fr_scratch_sizeof(n, w, b) == 32 * n + ba(w, b)
    assert orc.batch_affine_fr_scratch_sizeof(1, 10, 128) == 32
    for bad in ((0, 10, 128), (5, 0, 128), (5, 25, 128), (5, 10, 256), (1 << 32, 10, 128)):
        assert orc.batch_affine_fr_scratch_sizeof(*bad) == 0
    # the selected-algorithm query: the larger of the two plans inside
    # [2^12, 2^18], Pippenger outside, and past 2^18 never below the 2^18 plan
    assert orc.msm_fr_scratch_sizeof(4095) == orc.pippenger_fr_scratch_sizeof(4095)
    assert orc.msm_fr_scratch_sizeof(4096) == max(orc.pippenger_fr_scratch_sizeof(4096), 32 * 4096 + ba(10, 128))
    top = 32 * (1 << 18) + ba(15, 384)
    assert orc.msm_fr_scratch_sizeof(1 << 18) == max(orc.pippenger_fr_scratch_sizeof(1 << 18), top)
    assert orc.msm_fr_scratch_sizeof((1 << 18) + 1) == max(orc.pippenger_fr_scratch_sizeof((1 << 18) + 1), top)
    # capacity is monotone in n (BindingHeader_tests.cpp:1190-1203), sampled
    prev = 0
    for n in list(range(1, 5000)) + list(range(5000, (1 << 18) + 2, 97)) + [(1 << 18) + 1, 1 << 20]:
        cur = orc.msm_fr_scratch_sizeof(n)
        assert cur >= prev and cur >= orc.pippenger_fr_scratch_sizeof(n), n
        prev = cur


def test_g2_fr_scratch_sizes(orc):
    """The G2 blst_fr sizes (c_api.cpp:1976-2027 over the templates of
    :272-299, 593-627, 700-807): the 384-byte Pointxyzz<Fp2> buckets of the
    plain G2 entry (twice the G1 buckets pinned above), the batch-affine
    element sizes C's layout rules give the Fp2 structs of
    multi_scalar_batch_affine.cpp:16-36 (computed by ctypes), and the G2
    planner's batch-affine range of 2^12 .. 2^14 points."""
    import ctypes
    assert orc.p2_pippenger_fr_scratch_sizeof(0) == 0
    assert orc.p2_pippenger_fr_scratch_sizeof(1) == 32
    assert orc.p2_msm_fr_scratch_sizeof(0) == 0
    assert orc.p2_msm_fr_scratch_sizeof(1) == 32
    for n in (2, 3, 5, 64, 4096, 1 << 16, 1 << 20):
        assert orc.p2_pippenger_fr_scratch_sizeof(n) == 32 * n + 2 * orc.scratch_sizeof(n), n

    class AffBucket(ctypes.Structure):
        _fields_ = [("x", ctypes.c_uint64 * 12), ("y", ctypes.c_uint64 * 12),
                    ("generation", ctypes.c_uint32), ("queued_epoch", ctypes.c_uint32)]

    class ProjBucket(ctypes.Structure):
        _fields_ = [("point", ctypes.c_uint64 * 36), ("generation", ctypes.c_uint32)]

    assert (ctypes.sizeof(AffBucket), ctypes.sizeof(ProjBucket)) == (200, 296)

    def align(v, a):
        return -(-v // a) * a

    def ba(window, batch):
        B = 1 << (window - 1)
        off = B * ctypes.sizeof(AffBucket)
        off = align(off, ctypes.alignment(ProjBucket)) + B * ctypes.sizeof(ProjBucket)
        off = align(off, 4) + batch * 12
        off = align(off, 8) + batch * 96
        off = align(off, 8) + batch * 96
        return align(off, 8)

    for n, w, b in ((4096, 10, 128), (2, 1, 128), (5000, 11, 384), (7, 24, 128)):
        assert orc.p2_batch_affine_fr_scratch_sizeof(n, w, b) == 32 * n + ba(w, b)
    for bad in ((0, 10, 128), (5, 0, 128), (5, 25, 128), (5, 10, 256)):
        assert orc.p2_batch_affine_fr_scratch_sizeof(*bad) == 0
    assert orc.p2_msm_fr_scratch_sizeof(4095) == orc.p2_pippenger_fr_scratch_sizeof(4095)
    assert orc.p2_msm_fr_scratch_sizeof(4096) == max(orc.p2_pippenger_fr_scratch_sizeof(4096), 32 * 4096 + ba(10, 128))
    top = 32 * (1 << 14) + ba(11, 128)
    assert orc.p2_msm_fr_scratch_sizeof(1 << 14) == max(orc.p2_pippenger_fr_scratch_sizeof(1 << 14), top)
    # past 2^14 the plan is Pippenger, the capacity never below the 2^14 plan
    assert orc.p2_msm_fr_scratch_sizeof((1 << 14) + 1) == max(orc.p2_pippenger_fr_scratch_sizeof((1 << 14) + 1), top)
    prev = 0
    for n in list(range(1, 5000)) + list(range(5000, (1 << 18), 97)) + [1 << 20]:
        cur = orc.p2_msm_fr_scratch_sizeof(n)
        assert cur >= prev and cur >= orc.p2_pippenger_fr_scratch_sizeof(n), n
        prev = cur


def test_batch_to_affine_against_textbook(orc):
    """points_to_affine (multi_scalar.cpp:160-238) against x = X/Z^2,
    y = Y/Z^3 in Python big integers, with Z = 0 points (all-zero output)
    interleaved; and the scaled-Jacobian test-input helper round-trips."""
    rnd = random.Random(41)
    g = orc.from_affine(orc.generator_affine())
    jac = []
    for i in range(40):
        if i % 7 == 3:
            jac.append(np.zeros(18, np.uint64))
        else:
            jac.append(orc.p1_mult(g, rnd.getrandbits(255).to_bytes(32, "little"), 255))
    jac = np.stack(jac)
    aff = orc.p1s_to_affine(jac)
    for j, a in zip(jac, aff):
        want = pyref.jacobian_limbs_to_affine(list(j))
        if want is None:
            assert not a.any()
        else:
            assert pyref.limbs_to_affine(list(a)) == want
            assert np.array_equal(a, orc.to_affine(j))
    lam = np.stack([mont(rnd.randrange(1, P)) for _ in range(40)])
    back = orc.p1s_to_affine(orc.p1s_from_affine_scaled(aff, lam))
    assert np.array_equal(back, aff)
    for j, l in zip(orc.p1s_from_affine_scaled(aff, lam), lam):
        if j[12:].any():
            assert np.array_equal(j[12:], l)


def test_wbits_table_rows_against_textbook(orc):
    """Rows of point_precompute_row (multi_scalar.cpp:268-286) hold
    [1]P ... [2^(wbits-1)]P in affine; an all-zero point gives all-zero rows."""
    pts = orc.gen_points(4, 3, threads=1)
    pts[2] = 0
    for wb in (2, 3, 4, 6):
        tab = orc.wbits_precompute(pts, wb)
        m = 1 << (wb - 1)
        for i in range(4):
            P = pyref.limbs_to_affine(list(pts[i])) if pts[i].any() else None
            for j in range(m):
                e = tab[i * m + j]
                if P is None:
                    assert not e.any()
                else:
                    assert pyref.limbs_to_affine(list(e)) == pyref.mul(j + 1, P)


def test_wbits_msm_known_answers_and_textbook(orc):
    """C_api_tests.cpp:1606-1690: 2*[3]G + 3*[5]G + 5*[7]G at wbits 4 and
    the five-point case at wbits 10, nbits 8; then textbook sums over nbits
    and wbits, and agreement with the oracle's Pippenger."""
    def mults(ks):
        return np.stack([orc.to_affine(orc.p1_mult(orc.from_affine(orc.generator_affine()),
                                                    int(k).to_bytes(32, "little"), 255)) for k in ks])
    for ks, sc, wb, total in (([3, 5, 7], [2, 3, 5], 4, 56), ([3, 5, 7, 9, 11], [2, 3, 5, 7, 11], 10, 240)):
        pts = mults(ks)
        r = orc.mult_wbits(orc.wbits_precompute(pts, wb), wb, np.array(sc, np.uint8).reshape(-1, 1), len(ks), 8)
        assert orc.compress(r) == pyref.compress(pyref.mul(total, pyref.G))
    rnd = random.Random(17)
    pts = orc.gen_points(6, 19, threads=1)
    pts[4] = 0
    for nbits in (1, 2, 5, 8, 13, 64, 255, 256):
        nb = (nbits + 7) // 8
        ks = [rnd.getrandbits(8 * nb) for _ in range(6)]
        ks[1] = 0
        S = np.array([list(k.to_bytes(nb, "little")) for k in ks], np.uint8)
        for wb in (2, 3, 5, 8, 14):
            if wb == 14 and nbits < 64:
                continue
            r = orc.mult_wbits(orc.wbits_precompute(pts, wb), wb, S, 6, nbits)
            want = None
            for k, p in zip(ks, pts):
                if p.any():
                    want = pyref.add(want, pyref.mul(k % (1 << nbits), pyref.limbs_to_affine(list(p))))
            assert orc.compress(r) == pyref.compress(want), (nbits, wb)
            assert orc.compress(r) == orc.compress(orc.msm_pippenger(pts, S, 6, nbits))


def test_wbits_sizes(orc):
    """C_api_tests.cpp:2871-2873 and the closed forms of
    multi_scalar.h:244-256, 310-315."""
    assert orc.wbits_precompute_sizeof(0, 4) == 0
    assert orc.wbits_precompute_sizeof(15, 4) == 0
    assert orc.wbits_precompute_sizeof(14, (1 << 64) - 1) == 0
    assert orc.wbits_precompute_sizeof(1, 4) == 0
    assert orc.wbits_precompute_sizeof(2, 4) == 96 * 4 * 2
    assert orc.wbits_precompute_sizeof(14, 4096) == 96 * 4096 * 8192
    assert orc.wbits_scratch_sizeof(0) == 0 and orc.wbits_scratch_sizeof(3) == 3 * 144
    assert orc.wbits_scratch_sizeof(8192) == orc.wbits_scratch_sizeof(1 << 20) == 8192 * 144


# ---------------------------------------------------------------- G2

import pyref2  # noqa: E402


def g2_limbs(pt):
    return np.array(pyref2.affine_to_limbs(pt), dtype=np.uint64)


def test_fp2_mul_against_bigint(orc):
    rnd = random.Random(71)
    for _ in range(50):
        a = (rnd.randrange(P), rnd.randrange(P))
        b = (rnd.randrange(P), rnd.randrange(P))
        la = np.array(pyref.fp_limbs(pyref.to_mont(a[0])) + pyref.fp_limbs(pyref.to_mont(a[1])), np.uint64)
        lb = np.array(pyref.fp_limbs(pyref.to_mont(b[0])) + pyref.fp_limbs(pyref.to_mont(b[1])), np.uint64)
        r = orc.fp2_mul(la, lb)
        got = (pyref.from_mont(pyref.limbs_value(r[:6])), pyref.from_mont(pyref.limbs_value(r[6:])))
        assert got == pyref2.f2mul(a, b)


def test_g2_generator_and_hash_to_curve_points(orc):
    """The G2 generator (point.cpp:114-138) against the IETF encoding of
    Conformance_tests.cpp:18-21, and the 10 G2 hash-to-curve outputs."""
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    g = orc.p2_generator_affine()
    assert pyref2.limbs_to_affine(list(g)) == pyref2.G2
    assert orc.p2_compress(orc.p2_from_affine(g)).hex() == kat["g2_generator_compressed"]
    assert orc.p2_compress(np.zeros(36, np.uint64)).hex() == kat["g2_infinity_compressed"]
    for v in json.load(open(os.path.join(GOLD, "h2c_g2.json"))):
        pt = (tuple(int(c, 16) for c in v["x"]), tuple(int(c, 16) for c in v["y"]))
        assert pyref2.on_curve(pt)
        j = orc.p2_from_affine(g2_limbs(pt))
        assert orc.p2_compress(j).hex() == v["compressed"]
        assert orc.p2_compress(orc.p2_mult(j, pyref.R_ORDER.to_bytes(32, "little"), 256)) == \
            bytes([0xC0]) + bytes(95)      # in the prime-order subgroup


def test_g2_group_law_and_normalisation(orc):
    rnd = random.Random(72)
    G = orc.p2_from_affine(orc.p2_generator_affine())
    ks = [rnd.getrandbits(255) for _ in range(4)] + [1, 2, pyref.R_ORDER - 1]
    jac = np.stack([orc.p2_mult(G, k.to_bytes(32, "little"), 255) for k in ks] + [np.zeros(36, np.uint64)])
    want = [pyref2.mul(k, pyref2.G2) for k in ks] + [None]
    for j, w in zip(jac, want):
        assert pyref2.jacobian_limbs_to_affine(list(j)) == w
        assert orc.p2_compress(j) == pyref2.compress(w)
    aff = orc.p2s_to_affine(jac)
    assert [pyref2.limbs_to_affine(list(a)) for a in aff] == want
    s = orc.p2_add(jac[0], jac[1])
    assert orc.p2_compress(s) == pyref2.compress(pyref2.add(want[0], want[1]))
    assert orc.p2_compress(orc.p2_add(jac[0], jac[0])) == pyref2.compress(pyref2.add(want[0], want[0]))
    neg = orc.p2_mult(jac[0], (pyref.R_ORDER - 1).to_bytes(32, "little"), 255)
    assert orc.p2_compress(orc.p2_add(jac[0], neg)) == bytes([0xC0]) + bytes(95)


@pytest.mark.parametrize("n,nbits", [(1, 255), (2, 8), (5, 1), (7, 64), (9, 255), (12, 256)])
def test_g2_msm_against_textbook(orc, n, nbits):
    rnd = random.Random(n * 31 + nbits)
    pts = orc.gen_p2_points(n, 100 + n)
    if n > 3:
        pts[2] = 0
        pts[3] = orc.p2_to_affine(orc.p2_mult(orc.p2_from_affine(pts[1]), (pyref.R_ORDER - 1).to_bytes(32, "little"),
                                              255))
    nb = (nbits + 7) // 8
    ks = [rnd.getrandbits(8 * nb) for _ in range(n)]
    if n > 4:
        ks[4] = 0
        ks[3] = ks[1] % (1 << nbits)   # P and -P with equal scalars cancel
    S = np.array([list(k.to_bytes(nb, "little")) for k in ks], np.uint8)
    want = None
    for k, p in zip(ks, pts):
        if p.any():
            want = pyref2.add(want, pyref2.mul(k % (1 << nbits), pyref2.limbs_to_affine(list(p))))
    for window in (0, 1, 2, 3, 5, 8):
        assert orc.p2_compress(orc.msm_p2_pippenger(pts, S, n, nbits, window)) == pyref2.compress(want)
    # a single window's tile: the digits of bits [bit0 - 1, bit0 + 4)
    if nbits >= 9:
        t = orc.msm_p2_tile(pts, S, n, nbits, 4, 4)
        wt = None
        for k, p in zip(ks, pts):
            kk = k % (1 << nbits)
            d = ((kk >> 3) & 0x1F)
            dig = ((d + 1) >> 1) - (16 if d >> 4 else 0)
            if p.any() and dig:
                wt = pyref2.add(wt, pyref2.mul(dig % pyref.R_ORDER, pyref2.limbs_to_affine(list(p))))
        assert orc.p2_compress(t) == pyref2.compress(wt)


def test_table_window_digits_reconstruct_the_scalar(orc):
    """The digits of msm_mult_table's windows (or_table_booth: the top window
    holds nbits % wbits bits) recombine to the scalar itself:
    sum_w d_w 2^(rem_w) = k for every nbits and wbits."""
    rnd = random.Random(23)
    for nbits in (1, 2, 7, 8, 13, 64, 129, 255, 256):
        nb = (nbits + 7) // 8
        for wbits in (2, 3, 4, 7, 13, 14):
            for k in [0, 1, (1 << nbits) - 1, 1 << (nbits - 1)] + [rnd.getrandbits(nbits) for _ in range(20)]:
                s = np.array(list(k.to_bytes(nb, "little")), np.uint8)
                total, rem = 0, nbits - nbits % wbits
                while True:
                    e = orc.table_booth(s, nbits, wbits, rem)
                    mag = e & ((1 << wbits) - 1)
                    assert mag <= 1 << (wbits - 1)
                    total += (-mag if (e >> wbits) & 1 else mag) << rem
                    if rem == 0:
                        break
                    rem -= wbits
                assert total == k, (nbits, wbits, k)


def test_g2_wbits_tables_and_msm(orc):
    """G2 table rows hold [1]P ... [2^(wbits-1)]P (against the oracle's
    doubling-and-add scalar multiplication, pinned above against the textbook
    Fp2 model), an all-zero point gives all-zero rows, and the table MSM
    equals the G2 Pippenger MSM (pinned above) over nbits and wbits, with
    zero scalars, infinity, P and -P."""
    pts = orc.gen_p2_points(5, 7)
    pts[2] = 0
    pts[3] = orc.p2_to_affine(orc.p2_mult(orc.p2_from_affine(pts[1]), (pyref.R_ORDER - 1).to_bytes(32, "little"),
                                          255))
    for wb in (2, 3, 5):
        tab = orc.p2_wbits_precompute(pts, wb)
        m = 1 << (wb - 1)
        for i in range(5):
            for j in range(m):
                e = tab[i * m + j]
                if i == 2:
                    assert not e.any()
                else:
                    want = orc.p2_to_affine(orc.p2_mult(orc.p2_from_affine(pts[i]), (j + 1).to_bytes(2, "little"), 16))
                    assert np.array_equal(e, want), (wb, i, j)
    rnd = random.Random(29)
    for nbits in (1, 5, 8, 64, 255, 256):
        nb = (nbits + 7) // 8
        ks = [rnd.getrandbits(8 * nb) for _ in range(5)]
        ks[0] = 0
        ks[3] = ks[1]
        S = np.array([list(k.to_bytes(nb, "little")) for k in ks], np.uint8)
        want = orc.p2_compress(orc.msm_p2_pippenger(pts, S, 5, nbits))
        for wb in (2, 4, 7, 10):
            if wb == 10 and nbits < 64:
                continue
            r = orc.p2_mult_wbits(orc.p2_wbits_precompute(pts, wb), wb, S, 5, nbits)
            assert orc.p2_compress(r) == want, (nbits, wb)
