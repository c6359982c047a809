"""Independent pure-Python model of BLS12-381 G1 used to pin the oracle.

Field elements are Python integers mod p; points are affine (x, y) tuples
with None for infinity; the group law is the textbook short-Weierstrass
chord-and-tangent rule for y^2 = x^3 + 4.  Nothing here shares code with
the oracle or the product, so a wrong limb, sign or formula in either one
shows up as a mismatch.  Small cases only: every operation costs a
Python-level modular inversion.
"""

P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
R_ORDER = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
MONT_R = 1 << 384
GX = 0x17F1D3A73197D7942695638C4FA9AC0FC3688C4F9774B905A14E3A3F171BAC586C55E83FF97A1AEFFB3AF00ADB22C6BB
GY = 0x08B3F481E3AAA0F1A09E30ED741D8AE4FCF5E095D5D00AF600DB18CB2C04B3EDD03CC744A2888AE40CAA232946C5E7E1
G = (GX, GY)


def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return (y * y - x * x * x - 4) % P == 0


def neg(pt):
    if pt is None:
        return None
    return (pt[0], (-pt[1]) % P)


def add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    x1, y1 = a
    x2, y2 = b
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    y3 = (lam * (x1 - x3) - y1) % P
    return (x3, y3)


def mul(k, pt):
    acc = None
    base = pt
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
    """ZCash BLS12-381 compressed encoding (48 bytes)."""
    if pt is None:
        return bytes([0xC0]) + bytes(47)
    x, y = pt
    out = bytearray(x.to_bytes(48, "big"))
    out[0] |= 0x80
    if y > (P - 1) // 2:
        out[0] |= 0x20
    return bytes(out)


def decompress(data):
    data = bytes(data)
    assert data[0] & 0x80
    if data[0] & 0x40:
        return None
    big = bool(data[0] & 0x20)
    x = int.from_bytes(bytes([data[0] & 0x1F]) + data[1:], "big")
    y2 = (x * x * x + 4) % P
    y = pow(y2, (P + 1) // 4, P)
    assert y * y % P == y2
    if (y > (P - 1) // 2) != big:
        y = P - y
    return (x, y)


# --- conversions to/from the limb images used by the C ABI ------------------

def to_mont(x):
    return x * MONT_R % P


def from_mont(x):
    return x * pow(MONT_R, -1, P) % P


def fp_limbs(x):
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(6)]


def limbs_value(ls):
    return sum(int(v) << (64 * i) for i, v in enumerate(ls))


def affine_to_limbs(pt):
    """affine point -> 12 uint64 limbs (Montgomery x, y); infinity -> zeros"""
    if pt is None:
        return [0] * 12
    return fp_limbs(to_mont(pt[0])) + fp_limbs(to_mont(pt[1]))


def limbs_to_affine(ls):
    x = limbs_value(ls[:6])
    y = limbs_value(ls[6:12])
    if x == 0 and y == 0:
        return None
    return (from_mont(x), from_mont(y))


def jacobian_limbs_to_affine(ls):
    X = from_mont(limbs_value(ls[:6]))
    Y = from_mont(limbs_value(ls[6:12]))
    Z = from_mont(limbs_value(ls[12:18]))
    if Z == 0:
        return None
    zi = pow(Z, -1, P)
    return (X * zi * zi % P, Y * zi * zi * zi % P)
