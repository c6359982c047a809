"""Device field layer (fp.hip.h) against exact big-integer arithmetic.

The device keeps field elements in [0, 2p) (products without the final
subtraction, sums and differences reduced modulo 2p, canonical only on the
way out), so each operation is checked twice: its value modulo p, and the
range its output must stay in for the next operation's bound to hold.
Operands are 10^6 random words drawn uniformly from [0, 2p) and the boundary
values 0, 1, p - 1, p, p + 1, 2p - 1, (p +- 1)/2, R mod p and others, every
boundary value paired with every other as both operands.  The reference
statements are Fp.h:290-361 (mul, sqr, add, sub, cneg) and
backend_portable.cpp:19-175 (mont mul, add/sub mod p)."""
import random

import numpy as np
import pytest

import pyref

pytestmark = pytest.mark.gpu
P = pyref.P
R = 1 << 384
RINV = pow(R, -1, P)
N_RANDOM = 1 << 20

EDGES = sorted({0, 1, 2, 3, P - 2, P - 1, P, P + 1, P + 2, 2 * P - 2, 2 * P - 1, (P - 1) // 2, (P + 1) // 2,
                P + (P - 1) // 2, R % P, R % P + P, R * R % P, (R * R % P) + P, (1 << 380), (1 << 381) - 1,
                (1 << 32) - 1, (1 << 64) - 1, (1 << 352) - 1, (1 << 352), P - (1 << 32), P + (1 << 32),
                2 * P - (1 << 64), (2 * P) >> 1, (1 << 381) + 12345})
assert all(0 <= e < 2 * P for e in EDGES)


def _words(vals):
    return np.frombuffer(b"".join(v.to_bytes(48, "little") for v in vals), dtype=np.uint32).reshape(-1, 12)


def _ints(arr):
    b = np.ascontiguousarray(arr, dtype=np.uint32).tobytes()
    return [int.from_bytes(b[48 * i:48 * i + 48], "little") for i in range(len(b) // 48)]


@pytest.fixture(scope="module")
def operands():
    rnd = random.Random(0xF9)
    a = [rnd.randrange(2 * P) for _ in range(N_RANDOM)]
    b = [rnd.randrange(2 * P) for _ in range(N_RANDOM)]
    ea = [x for x in EDGES for _ in EDGES]
    eb = [y for _ in EDGES for y in EDGES]
    # every boundary value against random partners as well
    ra = [rnd.randrange(2 * P) for _ in range(len(EDGES) * 64)]
    a = ea + [x for x in EDGES for _ in range(64)] + ra + a
    b = eb + ra + [x for x in EDGES for _ in range(64)] + b
    return a, b, _words(a), _words(b)


def _check(got, want_mod_p, bound, what):
    bad = [k for k, (g, w) in enumerate(zip(got, want_mod_p)) if g >= bound or g % P != w]
    assert not bad, "%s: %d wrong, first at %d: got %x want %x mod p" % (
        what, len(bad), bad[0], got[bad[0]], want_mod_p[bad[0]])


def test_mul_and_sqr(hip, operands):
    a, b, A, B = operands
    got = _ints(hip.fp_ops("mul", A, B))
    _check(got, [x * y * RINV % P for x, y in zip(a, b)], 2 * P, "fp_mul")
    got = _ints(hip.fp_ops("sqr", A))
    _check(got, [x * x * RINV % P for x in a], 2 * P, "fp_sqr")


def test_mul_sum_one_reduction(hip, operands):
    """(a*b + c*d)/R in one reduction: below 8p^2/R + p < 1.82p for inputs
    below 2p (the bound the point formulas rely on); c runs over the
    negations 2p - y the adders feed it."""
    a, b, A, B = operands
    n = len(a)
    c = [(2 * P - y) if y else 0 for y in b[::-1]]
    d = a[::-1]
    got = _ints(hip.fp_ops("mul_sum", A, B, _words(c), _words(d)))
    _check(got, [(x * y + z * w) * RINV % P for x, y, z, w in zip(a, b, c, d)], (182 * P) // 100, "fp_mul_sum")
    assert len(got) == n


def test_add_sub_dbl(hip, operands):
    a, b, A, B = operands
    _check(_ints(hip.fp_ops("add", A, B)), [(x + y) % P for x, y in zip(a, b)], 2 * P, "fp_add")
    _check(_ints(hip.fp_ops("sub", A, B)), [(x - y) % P for x, y in zip(a, b)], 2 * P, "fp_sub")
    _check(_ints(hip.fp_ops("dbl", A)), [2 * x % P for x in a], 2 * P, "fp_dbl")


def test_neg_cneg_canon_is_zero(hip, operands):
    a, b, A, B = operands
    _check(_ints(hip.fp_ops("neg", A)), [(-x) % P for x in a], 2 * P, "fp_neg")
    # fp_neg maps 0 to 0 and p to p (both zero): never to 2p
    got = _ints(hip.fp_ops("neg", _words([0, P])))
    assert got == [0, P]
    _check(_ints(hip.fp_ops("cneg", A, B)), [(-x) % P if y & 1 else x % P for x, y in zip(a, b)], 2 * P,
           "fp_cneg")
    canon = _ints(hip.fp_ops("canon", A))
    assert canon == [x % P for x in a]
    z = _ints(hip.fp_ops("is_zero", A))
    assert z == [1 if x % P == 0 else 0 for x in a]


def test_inverse(hip, operands):
    """a^-1 in Montgomery form: the word image a = xR has inverse x^-1 R =
    R^2 / a; zero (0 or p) maps to zero."""
    a, _, _, _ = operands
    k = len(EDGES) ** 2 + (1 << 16)
    vals = a[:k]
    got = _ints(hip.fp_ops("inv", _words(vals)))
    want = [(R * R * pow(x, -1, P)) % P if x % P else 0 for x in vals]
    _check(got, want, 2 * P, "fp_inv")
