"""Independent pure-Python model of BLS12-381 G2 used to pin the oracle.

Fp2 = Fp[i] / (i^2 + 1), elements as (c0, c1) integer pairs; points are
affine ((x0, x1), (y0, y1)) with None for infinity on y^2 = x^3 + 4(1 + i),
with the textbook chord-and-tangent group law.  Nothing here shares code
with the oracle or the product.  Small cases only.
"""
from pyref import P, R_ORDER, MONT_R, fp_limbs, limbs_value

B2 = (4, 4)
G2 = ((0x024AA2B2F08F0A91260805272DC51051C6E47AD4FA403B02B4510B647AE3D1770BAC0326A805BBEFD48056C8C121BDB8,
       0x13E02B6052719F607DACD3A088274F65596BD0D09920B61AB5DA61BBDC7F5049334CF11213945D57E5AC7D055D042B7E),
      (0x0CE5D527727D6E118CC9CDC6DA2E351AADFD9BAA8CBDD3A76D429A695160D12C923AC9CC3BACA289E193548608B82801,
       0x0606C4A02EA734CC32ACD2B02BC28B99CB3E287E85A763AF267492AB572E99AB3F370D275CEC1DA1AAA9075FF05F79BE))


def f2add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2inv(a):
    n = pow(a[0] * a[0] + a[1] * a[1], -1, P)
    return (a[0] * n % P, (-a[1]) * n % P)


def f2neg(a):
    return ((-a[0]) % P, (-a[1]) % P)


def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return f2sub(f2mul(y, y), f2add(f2mul(f2mul(x, x), x), B2)) == (0, 0)


def neg(pt):
    return None if pt is None else (pt[0], f2neg(pt[1]))


def add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    (x1, y1), (x2, y2) = a, b
    if x1 == x2:
        if f2add(y1, y2) == (0, 0):
            return None
        lam = f2mul(f2mul((3, 0), f2mul(x1, x1)), f2inv(f2add(y1, y1)))
    else:
        lam = f2mul(f2sub(y2, y1), f2inv(f2sub(x2, x1)))
    x3 = f2sub(f2sub(f2mul(lam, lam), x1), x2)
    y3 = f2sub(f2mul(lam, f2sub(x1, x3)), y1)
    return (x3, y3)


def mul(k, pt):
    acc, base = None, pt
    while k > 0:
        if k & 1:
            acc = add(acc, base)
        base = add(base, base)
        k >>= 1
    return acc


def msm(points, scalars):
    acc = None
    for p, k in zip(points, scalars):
        acc = add(acc, mul(k, p))
    return acc


def compress(pt):
    """ZCash BLS12-381 G2 compressed encoding (96 bytes: x.c1 then x.c0,
    big-endian; the sign flag is that of y.c1, or of y.c0 when y.c1 = 0)."""
    if pt is None:
        return bytes([0xC0]) + bytes(95)
    (x0, x1), (y0, y1) = pt
    out = bytearray(x1.to_bytes(48, "big") + x0.to_bytes(48, "big"))
    out[0] |= 0x80
    big = y1 > (P - 1) // 2 if y1 else y0 > (P - 1) // 2
    if big:
        out[0] |= 0x20
    return bytes(out)


def to_mont(x):
    return x * MONT_R % P


def from_mont(x):
    return x * pow(MONT_R, -1, P) % P


def affine_to_limbs(pt):
    """G2 affine -> 24 uint64 limbs (x.c0, x.c1, y.c0, y.c1, Montgomery)"""
    if pt is None:
        return [0] * 24
    (x0, x1), (y0, y1) = pt
    return sum((fp_limbs(to_mont(v)) for v in (x0, x1, y0, y1)), [])


def limbs_to_affine(ls):
    vals = [limbs_value(ls[6 * k:6 * k + 6]) for k in range(4)]
    if not any(vals):
        return None
    x0, x1, y0, y1 = (from_mont(v) for v in vals)
    return ((x0, x1), (y0, y1))


def jacobian_limbs_to_affine(ls):
    X, Y, Z = [(from_mont(limbs_value(ls[12 * k:12 * k + 6])), from_mont(limbs_value(ls[12 * k + 6:12 * k + 12])))
               for k in range(3)]
    if Z == (0, 0):
        return None
    zi = f2inv(Z)
    zi2 = f2mul(zi, zi)
    return (f2mul(X, zi2), f2mul(Y, f2mul(zi2, zi)))


__all__ = ["P", "R_ORDER", "G2"]
