"""The 28-bit limb field of the G1 bucket accumulation
(temp-cat-blst_amd/csrc/fq.hip.h) against Python big integers, at random
inputs and at the largest limbs and values each operation admits, and its
bucket sum against the textbook group law (tests/pyref.py): doublings,
P + (-P), negated and infinite inputs, and sums continued from a stored one.

The same header is checked twice: compiled by g++ for the host (CPU suite)
and on the device through the test kit's blst_hip_fq_ops (GPU suite)."""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

import pyref

HERE = os.path.dirname(os.path.abspath(__file__))
P = pyref.P
W, N = 28, 14
MASK = (1 << W) - 1
RINV = pow(1 << 392, -1, P)
OPS = {"mul": 0, "sqr": 1, "mul_sum": 2, "reduce": 3, "sub_lazy": 4, "neg_lazy": 5, "x3": 6, "is_zero": 7,
       "from_words8": 8, "from_words4": 9, "from_words0": 10, "to_words": 11, "times4": 12,
       "reduce_wide": 13, "canon": 14}


def limbs(x):
    return [(x >> (W * i)) & MASK for i in range(N)]


def val(row):
    return sum(int(v) << (W * i) for i, v in enumerate(row))


def borrow_form(v, b):
    """v as 14 limbs with b 2^28 moved into every limb below the top one"""
    n = limbs(v)
    return [n[i] + ((b << W) if i < N - 1 else 0) - (b if i > 0 else 0) for i in range(N)]


K4, K6 = borrow_form(4 * P, 1), borrow_form(6 * P, 3)
assert val(K4) == 4 * P and val(K6) == 6 * P


def rand_limbs(rnd, maxlimb, maxval):
    while True:
        row = [rnd.randrange(maxlimb) for _ in range(N - 1)]
        row.append(rnd.randrange(1 << 21))
        if val(row) < maxval:
            return row


def top_limbs(maxlimb, maxval):
    """every limb at maxlimb - 1 with the top limb as large as maxval allows"""
    row = [maxlimb - 1] * (N - 1)
    return row + [max((maxval - 1 - val(row + [0])) >> (W * (N - 1)), 0)]


LAZY = int(2 ** 29.6)          # limbs of a - b + 4p for normalized a, b


def op_cases(rnd, count):
    """(op, a, b, c, d, check) tuples: inputs at the bounds each operation
    states, check(a, b, c, d, r) on the Python values"""
    cases = []

    def norm(maxval):
        return limbs(rnd.randrange(maxval))

    def normal(r, bound):
        return all(v <= MASK for v in r) and val(r) < bound

    def chk_mul(a, b, c, d, r):
        A, B = val(a), val(b)
        return val(r) % P == A * B * RINV % P and normal(r, A * B // (1 << 392) + P + 1)

    def chk_sqr(a, b, c, d, r):
        A = val(a)
        return val(r) % P == A * A * RINV % P and normal(r, A * A // (1 << 392) + P + 1)

    def chk_ms(a, b, c, d, r):
        S = val(a) * val(b) + val(c) * val(d)
        return val(r) % P == S * RINV % P and normal(r, S // (1 << 392) + P + 1)

    def chk_red(a, b, c, d, r):
        return val(r) % P == val(a) % P and normal(r, 2 * P)

    def chk_sub(a, b, c, d, r):
        return val(r) == val(a) - val(b) + 4 * P and all(v < LAZY for v in r)

    def chk_neg(a, b, c, d, r):
        return val(r) == 4 * P - val(a) and all(v < int(2 ** 29.1) for v in r)

    def chk_x3(a, b, c, d, r):
        s = (0 if d[0] & 1 else val(b)) + 2 * val(c)
        return val(r) % P == (val(a) - s) % P and normal(r, 2 * P)

    def chk_zero(a, b, c, d, r):
        return r[0] == (1 if val(a) % P == 0 else 0)

    def chk_wide(a, b, c, d, r):
        bound = P * 101 // 100 if val(a) < 1 << 390 else P * 103 // 100
        return val(r) % P == val(a) % P and normal(r, bound)

    def chk_canon(a, b, c, d, r):
        return [int(x) for x in r] == limbs(val(a) % P)

    for k in range(count):
        edge = k % 4 == 0
        # products: normalized, lazy, and the loaded x (below 2^390) against a value below 2p
        for A, B in ((norm(2 * P), norm(2 * P)),
                     (top_limbs(LAZY, 9 * P), top_limbs(LAZY, 9 * P)) if edge else
                     (rand_limbs(rnd, LAZY, 9 * P), rand_limbs(rnd, LAZY, 9 * P)),
                     (norm(1 << 390), norm(2 * P))):
            cases.append(("mul", A, B, A, A, chk_mul))
            cases.append(("sqr", A, A, A, A, chk_sqr))
        C = top_limbs(int(2 ** 29.1), 4 * P) if edge else rand_limbs(rnd, int(2 ** 29.1), 4 * P)
        Dn = top_limbs(1 << W, 2 * P) if edge else norm(2 * P)
        A = top_limbs(LAZY, 6 * P) if edge else rand_limbs(rnd, LAZY, 6 * P)
        cases.append(("mul_sum", A, A, C, Dn, chk_ms))
        X = top_limbs(int(2 ** 30.4), 8 * P) if edge else rand_limbs(rnd, int(2 ** 30.4), 8 * P)
        cases.append(("reduce", X, X, X, X, chk_red))
        a, b = norm(int(1.2 * P)), norm(2 * P)
        cases.append(("sub_lazy", a, b, a, a, chk_sub))
        cases.append(("neg_lazy", b, b, b, b, chk_neg))
        r2, ppp, q = (norm(int(1.1 * P)) for _ in range(3))
        if edge:
            r2, ppp, q = [0] * N, top_limbs(1 << W, int(1.1 * P)), top_limbs(1 << W, int(1.1 * P))
        dbl = [k & 1] + [0] * (N - 1)
        cases.append(("x3", r2, ppp, q, dbl, chk_x3))
        z = [0, P, 2 * P - 1, 1, rnd.randrange(2 * P)][k % 5]
        cases.append(("is_zero", limbs(z), limbs(z), limbs(z), limbs(z), chk_zero))
        # any normalized value: the largest, near multiples of p, the exact
        # images of 32-bit words (w 2^8, w below 2p) and random ones
        m = rnd.randrange(1, (1 << 392) // P)
        v = [(1 << 392) - 1, m * P - 1, m * P, m * P + 1, rnd.randrange(2 * P) << 8,
             rnd.randrange(1 << 392)][k % 6]
        cases.append(("reduce_wide", limbs(v), limbs(v), limbs(v), limbs(v), chk_wide))
        v = [0, 1, P - 1, P, P + 1, 2 * P - 1, rnd.randrange(2 * P)][k % 7]
        cases.append(("canon", limbs(v), limbs(v), limbs(v), limbs(v), chk_canon))
    return cases


def words_cases(rnd, count):
    out = []
    for k in range(count):
        v = [0, 1, P - 1, 2 * P - 1, rnd.randrange(2 * P)][k % 5]
        w = [(v >> (32 * i)) & 0xFFFFFFFF for i in range(12)] + [0, 0]
        out.append((v, w))
    return out


def run_op_checks(run, seed, count):
    """run(op, A, B, C, D) -> (n, 14) results for arrays of rows"""
    rnd = random.Random(seed)
    cases = op_cases(rnd, count)
    by_op = {}
    for c in cases:
        by_op.setdefault(c[0], []).append(c)
    checked = 0
    for op, cs in by_op.items():
        arrs = [np.array([c[j] for c in cs], np.uint32) for j in (1, 2, 3, 4)]
        res = run(op, *arrs)
        for c, r in zip(cs, res):
            assert c[5](c[1], c[2], c[3], c[4], [int(v) for v in r]), (op, c[1], c[2])
            checked += 1
    wc = words_cases(rnd, count)
    W_ = np.array([w for _, w in wc], np.uint32)
    for op, shift in (("from_words8", 8), ("from_words4", 4), ("from_words0", 0)):
        res = run(op, W_, W_, W_, W_)
        for (v, _), r in zip(wc, res):
            assert val(r) == v << shift and all(int(x) <= MASK for x in r), (op, v)
    L_ = np.array([limbs(v) for v, _ in wc], np.uint32)
    res = run("to_words", L_, L_, L_, L_)
    for (v, w), r in zip(wc, res):
        assert [int(x) for x in r[:12]] == w[:12]
    half = np.array([limbs(v % (2 * P)) for v, _ in wc], np.uint32)
    res = run("times4", half, half, half, half)
    for row, r in zip(half, res):
        assert val(r) % P == 4 * val(row) % P and val(r) < 2 * P
    return checked


@pytest.fixture(scope="module")
def host_fq(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fq") / "libfq_host.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas", "-o", out,
                           os.path.join(HERE, "fq_host.cpp")])
    L = ctypes.CDLL(out)
    vp = ctypes.c_void_p
    L.fqh_ops.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_size_t]
    L.fqh_accumulate.argtypes = [vp, vp, vp, ctypes.c_size_t]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_fq_ops_host(host_fq):
    def run(op, A, B, C, D):
        A, B, C, D = (np.ascontiguousarray(x, np.uint32) for x in (A, B, C, D))
        out = np.zeros_like(A)
        host_fq.fqh_ops(OPS[op], _p(out), _p(A), _p(B), _p(C), _p(D), A.shape[0])
        return out
    assert run_op_checks(run, 1, 400) > 4000


# ------------------------------------------------------------ bucket sums

def _words(x):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(12)]


def _wval(w):
    return sum(int(v) << (32 * i) for i, v in enumerate(w))


def _pt_words(pt):
    if pt is None:
        return [0] * 24
    return _words(pyref.to_mont(pt[0])) + _words(pyref.to_mont(pt[1]))


def _xyzz_affine(io):
    """the affine point of a 32-bit path XYZZ word image; checks the range and
    ZZ^3 = ZZZ^2"""
    X, Y, ZZ, ZZZ = (_wval(io[12 * c:12 * c + 12]) for c in range(4))
    assert all(v < 2 * P for v in (X, Y, ZZ, ZZZ))
    if ZZ % P == 0:
        return None
    ri = pow(1 << 384, -1, P)
    X, Y, ZZ, ZZZ = (v * ri % P for v in (X, Y, ZZ, ZZZ))
    assert pow(ZZ, 3, P) == pow(ZZZ, 2, P)
    return X * pow(ZZ, -1, P) % P, Y * pow(ZZZ, -1, P) % P


def _xyzz_words(pt, rnd):
    """a random XYZZ word image of pt (some coordinates in [p, 2p))"""
    if pt is None:
        return [0] * 48
    z = rnd.randrange(1, P)
    out = []
    for v in (pt[0] * z * z % P, pt[1] * z * z * z % P, z * z % P, z * z * z % P):
        m = pyref.to_mont(v)
        if rnd.random() < 0.3 and m + P < 2 * P:
            m += P
        out += _words(m)
    return out


def _accumulate(L, start, pts, negs):
    io = np.array(start, np.uint32)
    flat = np.array([w for p in pts for w in _pt_words(p)] or [0], np.uint32)
    ng = np.array(negs or [0], np.uint8)
    L.fqh_accumulate(_p(io), _p(flat), _p(ng), len(pts))
    return _xyzz_affine([int(v) for v in io])


def test_fq_bucket_sums_host(host_fq):
    rnd = random.Random(5)
    base = [pyref.mul(rnd.randrange(1, 1 << 64), pyref.G) for _ in range(12)]
    for trial in range(250):
        n = rnd.randrange(1, 12)
        pts = [rnd.choice(base) for _ in range(n)]
        negs = [int(rnd.random() < 0.4) for _ in range(n)]
        if rnd.random() < 0.3:
            pts[rnd.randrange(n)] = None
        if n >= 2 and rnd.random() < 0.4:   # the same point again: doubling or cancellation
            i = rnd.randrange(n - 1)
            pts[i + 1] = pts[i]
            negs[i + 1] = negs[i] if rnd.random() < 0.5 else 1 - negs[i]
        start = rnd.choice([None, None, rnd.choice(base)])
        want = start
        for p, s in zip(pts, negs):
            if p is not None:
                want = pyref.add(want, pyref.neg(p) if s else p)
        assert _accumulate(host_fq, _xyzz_words(start, rnd), pts, negs) == want, trial
    g = base[0]
    assert _accumulate(host_fq, [0] * 48, [g] * 40, [0] * 40) == pyref.mul(40, g)
    assert _accumulate(host_fq, [0] * 48, [g, g, g, g], [0, 1, 0, 1]) is None
    assert _accumulate(host_fq, [0] * 48, [g, g, base[1]], [0, 1, 0]) == base[1]
    assert _accumulate(host_fq, [0] * 48, [None, None], [0, 0]) is None


@pytest.mark.gpu
def test_fq_ops_device(hip):
    def run(op, A, B, C, D):
        return hip.fq_ops(op, A, B, C, D)
    print("fq ops on the device", flush=True)
    assert run_op_checks(run, 2, 2000) > 20000
