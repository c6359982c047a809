"""TEST INFRASTRUCTURE ONLY: ctypes binding of the oracle (liboracle.so).

The oracle is a CPU restatement of category-blst's G1 MSM path (see
blst_oracle.c for the file:line each function follows).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only to check results; the product library never touches it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_sz = ctypes.c_size_t

_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE, "-B"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        sigs = {
            "or_fp_mul": [_u64p, _u64p, _u64p],
            "or_fp_add": [_u64p, _u64p, _u64p],
            "or_fp_sub": [_u64p, _u64p, _u64p],
            "or_fp_neg": [_u64p, _u64p],
            "or_fp_to_mont": [_u64p, _u64p],
            "or_fp_from_mont": [_u64p, _u64p],
            "or_fp_inv": [_u64p, _u64p],
            "or_p1_generator": [_u64p],
            "or_p1_from_affine": [_u64p, _u64p],
            "or_p1_add": [_u64p, _u64p, _u64p],
            "or_p1_double": [_u64p, _u64p],
            "or_p1_mult": [_u64p, _u64p, _u8p, _sz],
            "or_p1_to_affine": [_u64p, _u64p],
            "or_p1_compress": [_u8p, _u64p],
            "or_p1s_to_affine": [_u64p, _u64p, _sz],
            "or_p1s_from_affine_scaled": [_u64p, _u64p, _u64p, _sz],
            "or_msm_pippenger": [_u64p, _u64p, _u8p, _sz, _sz, _sz],
            "or_msm_pippenger_mt": [_u64p, _u64p, _u8p, _sz, _sz, _sz, ctypes.c_int],
            "or_msm_tile": [_u64p, _u64p, _u8p, _sz, _sz, _sz, _sz],
            "or_msm_batch_mt": [_u64p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _sz,
                                ctypes.c_int],
            "or_msm_naive": [_u64p, _u64p, _u8p, _sz, _sz],
            "or_gen_scalars": [_u8p, _sz, ctypes.c_uint64, ctypes.c_int],
            "or_gen_points": [_u64p, _sz, ctypes.c_uint64, ctypes.c_int],
            "or_gen_srs": [_u64p, _sz, ctypes.c_uint64, ctypes.c_int],
            "or_apply_edge_cases": [_u64p, _u8p, _sz, ctypes.c_uint64],
            "or_fr_to_scalar": [_u8p, _u64p],
            "or_frs_to_scalars": [_u8p, _u64p, _sz],
            "or_wbits_precompute": [_u64p, _sz, _u64p, _sz],
            "or_fp2_mul": [_u64p, _u64p, _u64p],
            "or_p2_generator": [_u64p],
            "or_p2_from_affine": [_u64p, _u64p],
            "or_p2_add": [_u64p, _u64p, _u64p],
            "or_p2_double": [_u64p, _u64p],
            "or_p2_to_affine": [_u64p, _u64p],
            "or_p2_mult": [_u64p, _u64p, _u8p, _sz],
            "or_p2s_to_affine": [_u64p, _u64p, _sz],
            "or_p2_compress": [_u8p, _u64p],
            "or_msm_p2_tile": [_u64p, _u64p, _u8p, _sz, _sz, _sz, _sz],
            "or_msm_p2_pippenger": [_u64p, _u64p, _u8p, _sz, _sz, _sz],
            "or_gen_p2_points": [_u64p, _sz, ctypes.c_uint64],
            "or_mult_wbits": [_u64p, _u64p, _sz, _sz, _u8p, _sz],
            "or_p2_wbits_precompute": [_u64p, _sz, _u64p, _sz],
            "or_p2_mult_wbits": [_u64p, _u64p, _sz, _sz, _u8p, _sz],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        for name, args, res in [
            ("or_fp_sgn0_pty_mont", [_u64p], ctypes.c_uint64),
            ("or_p1_is_equal", [_u64p, _u64p], ctypes.c_int),
            ("or_p1_affine_on_curve", [_u64p], ctypes.c_int),
            ("or_pippenger_window_size", [_sz], _sz),
            ("or_pippenger_scratch_sizeof", [_sz], _sz),
            ("or_pippenger_window_scratch_sizeof", [_sz], _sz),
            ("or_booth_digit", [_u8p, _sz, _sz, _sz], ctypes.c_uint64),
            ("or_pippenger_fr_scratch_sizeof", [_sz], _sz),
            ("or_batch_affine_fr_scratch_sizeof", [_sz, _sz, _sz], _sz),
            ("or_msm_fr_scratch_sizeof", [_sz], _sz),
            ("or_p2_pippenger_fr_scratch_sizeof", [_sz], _sz),
            ("or_p2_batch_affine_fr_scratch_sizeof", [_sz, _sz, _sz], _sz),
            ("or_p2_msm_fr_scratch_sizeof", [_sz], _sz),
            ("or_wbits_precompute_sizeof", [_sz, _sz], _sz),
            ("or_wbits_scratch_sizeof", [_sz], _sz),
            ("or_table_booth", [_u8p, _sz, _sz, _sz], ctypes.c_uint64),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def u64p(a):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(_u64p)


def u8p(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_u8p)


# ---- field helpers on python ints (values, not Montgomery) -----------------

def limbs(x, n=6):
    return np.array([(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(n)], dtype=np.uint64)


def value(a):
    return sum(int(v) << (64 * i) for i, v in enumerate(a))


def fp_mul(a, b):
    r = np.zeros(6, np.uint64)
    lib().or_fp_mul(u64p(r), u64p(np.ascontiguousarray(a)), u64p(np.ascontiguousarray(b)))
    return r


def fp_binop(name, a, b):
    r = np.zeros(6, np.uint64)
    getattr(lib(), name)(u64p(r), u64p(np.ascontiguousarray(a)), u64p(np.ascontiguousarray(b)))
    return r


def fp_unop(name, a):
    r = np.zeros(6, np.uint64)
    getattr(lib(), name)(u64p(r), u64p(np.ascontiguousarray(a)))
    return r


# ---- points ----------------------------------------------------------------

def generator_affine():
    g = np.zeros(12, np.uint64)
    lib().or_p1_generator(u64p(g))
    return g


def from_affine(a):
    r = np.zeros(18, np.uint64)
    lib().or_p1_from_affine(u64p(r), u64p(np.ascontiguousarray(a, dtype=np.uint64)))
    return r


def p1_add(a, b):
    r = np.zeros(18, np.uint64)
    lib().or_p1_add(u64p(r), u64p(np.ascontiguousarray(a)), u64p(np.ascontiguousarray(b)))
    return r


def p1_double(a):
    r = np.zeros(18, np.uint64)
    lib().or_p1_double(u64p(r), u64p(np.ascontiguousarray(a)))
    return r


def p1_mult(p, scalar_bytes, nbits):
    r = np.zeros(18, np.uint64)
    s = np.frombuffer(bytes(scalar_bytes), dtype=np.uint8).copy()
    lib().or_p1_mult(u64p(r), u64p(np.ascontiguousarray(p)), u8p(s), nbits)
    return r


def to_affine(p):
    r = np.zeros(12, np.uint64)
    lib().or_p1_to_affine(u64p(r), u64p(np.ascontiguousarray(p)))
    return r


def p1s_to_affine(jac):
    """(n, 18) Jacobian -> (n, 12) affine (points_to_affine)."""
    jac = np.ascontiguousarray(jac, dtype=np.uint64)
    out = np.zeros((jac.shape[0], 12), np.uint64)
    if jac.shape[0]:
        lib().or_p1s_to_affine(u64p(out.reshape(-1)), u64p(jac.reshape(-1)), jac.shape[0])
    return out


def p1s_from_affine_scaled(aff, lam):
    """(n, 12) affine and (n, 6) Montgomery lambdas -> (n, 18) Jacobian
    (x l^2, y l^3, l); all-zero affine -> Z = 0."""
    aff = np.ascontiguousarray(aff, dtype=np.uint64)
    lam = np.ascontiguousarray(lam, dtype=np.uint64)
    out = np.zeros((aff.shape[0], 18), np.uint64)
    if aff.shape[0]:
        lib().or_p1s_from_affine_scaled(u64p(out.reshape(-1)), u64p(aff.reshape(-1)), u64p(lam.reshape(-1)),
                                        aff.shape[0])
    return out


def compress(p):
    out = np.zeros(48, np.uint8)
    lib().or_p1_compress(u8p(out), u64p(np.ascontiguousarray(p, dtype=np.uint64)))
    return out.tobytes()


def is_equal(a, b):
    return bool(lib().or_p1_is_equal(u64p(np.ascontiguousarray(a)), u64p(np.ascontiguousarray(b))))


def on_curve(a):
    return bool(lib().or_p1_affine_on_curve(u64p(np.ascontiguousarray(a, dtype=np.uint64))))


# ---- MSM -------------------------------------------------------------------

def window_size(n):
    return lib().or_pippenger_window_size(n)


def scratch_sizeof(n):
    return lib().or_pippenger_scratch_sizeof(n)


def window_scratch_sizeof(w):
    return lib().or_pippenger_window_scratch_sizeof(w)


def booth_digit(scalar_bytes, nbits, bit0, window):
    s = np.frombuffer(bytes(scalar_bytes), dtype=np.uint8).copy()
    return lib().or_booth_digit(u8p(s), nbits, bit0, window)


def msm_pippenger(points, scalars, n, nbits, window=0, threads=1):
    """points: (n,12) uint64 affine Montgomery; scalars: (n, nbytes) uint8."""
    r = np.zeros(18, np.uint64)
    pts = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
    if threads > 1:
        lib().or_msm_pippenger_mt(u64p(r), u64p(pts), u8p(sc), n, nbits, window, threads)
    else:
        lib().or_msm_pippenger(u64p(r), u64p(pts), u8p(sc), n, nbits, window)
    return r


def msm_batch(jobs, threads=16):
    """Independent MSMs on a thread pool.  jobs: (points, scalars, n, nbits)
    with points (>= n, 12) uint64 and scalars (>= n, nbytes) uint8 arrays;
    returns (njobs, 18) uint64 Jacobian results."""
    nj = len(jobs)
    keep = []
    pp = np.zeros(max(nj, 1), np.uint64)
    sp = np.zeros(max(nj, 1), np.uint64)
    ns = np.zeros(max(nj, 1), np.uint64)
    nb = np.zeros(max(nj, 1), np.uint64)
    for k, (pts, sc, n, nbits) in enumerate(jobs):
        P = np.ascontiguousarray(pts, dtype=np.uint64)
        S = np.ascontiguousarray(sc, dtype=np.uint8)
        keep += [P, S]
        pp[k], sp[k], ns[k], nb[k] = P.ctypes.data, S.ctypes.data, n, nbits
    out = np.zeros((max(nj, 1), 18), np.uint64)
    lib().or_msm_batch_mt(u64p(out.reshape(-1)), ctypes.c_void_p(pp.ctypes.data), ctypes.c_void_p(sp.ctypes.data),
                          ctypes.c_void_p(ns.ctypes.data), ctypes.c_void_p(nb.ctypes.data), nj, threads)
    return out[:nj]


def msm_tile(points, scalars, n, nbits, bit0, window):
    r = np.zeros(18, np.uint64)
    pts = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
    lib().or_msm_tile(u64p(r), u64p(pts), u8p(sc), n, nbits, bit0, window)
    return r


def msm_naive(points, scalars, n, nbits):
    r = np.zeros(18, np.uint64)
    pts = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
    lib().or_msm_naive(u64p(r), u64p(pts), u8p(sc), n, nbits)
    return r


# ---- blst_fr ---------------------------------------------------------------

R_ORDER = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001


def fr_to_scalars(fr):
    """(n, 4) uint64 Montgomery-form blst_fr -> (n, 32) uint8 canonical scalars."""
    fr = np.ascontiguousarray(fr, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros((fr.shape[0], 32), np.uint8)
    if fr.shape[0]:
        lib().or_frs_to_scalars(u8p(out.reshape(-1)), u64p(fr.reshape(-1)), fr.shape[0])
    return out


def fr_from_values(vals):
    """Python ints below r -> (n, 4) uint64 Montgomery form (x * 2^256 mod r)."""
    return np.stack([limbs((int(v) << 256) % R_ORDER, 4) for v in vals]) if len(vals) else \
        np.zeros((0, 4), np.uint64)


def pippenger_fr_scratch_sizeof(n):
    return lib().or_pippenger_fr_scratch_sizeof(n)


def batch_affine_fr_scratch_sizeof(n, window, batch):
    return lib().or_batch_affine_fr_scratch_sizeof(n, window, batch)


def msm_fr_scratch_sizeof(n):
    return lib().or_msm_fr_scratch_sizeof(n)


# ---- fixed-base tables -----------------------------------------------------

def p2_pippenger_fr_scratch_sizeof(n):
    return lib().or_p2_pippenger_fr_scratch_sizeof(n)


def p2_batch_affine_fr_scratch_sizeof(n, window, batch):
    return lib().or_p2_batch_affine_fr_scratch_sizeof(n, window, batch)


def p2_msm_fr_scratch_sizeof(n):
    return lib().or_p2_msm_fr_scratch_sizeof(n)


def wbits_precompute_sizeof(wbits, n):
    return lib().or_wbits_precompute_sizeof(wbits, n)


def wbits_scratch_sizeof(n):
    return lib().or_wbits_scratch_sizeof(n)


def wbits_precompute(points, wbits):
    """(n, 12) affine -> (n * 2^(wbits-1), 12) affine table, rows of [j]P."""
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    n = pts.shape[0]
    out = np.zeros((n << (wbits - 1), 12), np.uint64)
    if n:
        lib().or_wbits_precompute(u64p(out.reshape(-1)), wbits, u64p(pts.reshape(-1)), n)
    return out


def mult_wbits(table, wbits, scalars, n, nbits):
    r = np.zeros(18, np.uint64)
    tab = np.ascontiguousarray(table, dtype=np.uint64).reshape(-1)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
    lib().or_mult_wbits(u64p(r), u64p(tab), wbits, n, u8p(sc), nbits)
    return r


# ---- G2 ----------------------------------------------------------------------

def fp2_mul(a, b):
    r = np.zeros(12, np.uint64)
    lib().or_fp2_mul(u64p(r), u64p(np.ascontiguousarray(a, dtype=np.uint64)),
                     u64p(np.ascontiguousarray(b, dtype=np.uint64)))
    return r


def p2_generator_affine():
    g = np.zeros(24, np.uint64)
    lib().or_p2_generator(u64p(g))
    return g


def p2_from_affine(a):
    r = np.zeros(36, np.uint64)
    lib().or_p2_from_affine(u64p(r), u64p(np.ascontiguousarray(a, dtype=np.uint64)))
    return r


def p2_add(a, b):
    r = np.zeros(36, np.uint64)
    lib().or_p2_add(u64p(r), u64p(np.ascontiguousarray(a)), u64p(np.ascontiguousarray(b)))
    return r


def p2_mult(p, scalar_bytes, nbits):
    r = np.zeros(36, np.uint64)
    s = np.frombuffer(bytes(scalar_bytes), dtype=np.uint8).copy()
    lib().or_p2_mult(u64p(r), u64p(np.ascontiguousarray(p)), u8p(s), nbits)
    return r


def p2_to_affine(p):
    r = np.zeros(24, np.uint64)
    lib().or_p2_to_affine(u64p(r), u64p(np.ascontiguousarray(p)))
    return r


def p2s_to_affine(jac):
    jac = np.ascontiguousarray(jac, dtype=np.uint64)
    out = np.zeros((jac.shape[0], 24), np.uint64)
    if jac.shape[0]:
        lib().or_p2s_to_affine(u64p(out.reshape(-1)), u64p(jac.reshape(-1)), jac.shape[0])
    return out


def p2_compress(p):
    out = np.zeros(96, np.uint8)
    lib().or_p2_compress(u8p(out), u64p(np.ascontiguousarray(p, dtype=np.uint64)))
    return out.tobytes()


def msm_p2_pippenger(points, scalars, n, nbits, window=0):
    r = np.zeros(36, np.uint64)
    pts = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
    lib().or_msm_p2_pippenger(u64p(r), u64p(pts), u8p(sc), n, nbits, window)
    return r


def msm_p2_tile(points, scalars, n, nbits, bit0, window):
    r = np.zeros(36, np.uint64)
    pts = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
    lib().or_msm_p2_tile(u64p(r), u64p(pts), u8p(sc), n, nbits, bit0, window)
    return r


def p2_wbits_precompute(points, wbits):
    """(n, 24) G2 affine -> (n * 2^(wbits-1), 24) affine table, rows of [j]P."""
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    n = pts.shape[0]
    out = np.zeros((n << (wbits - 1), 24), np.uint64)
    if n:
        lib().or_p2_wbits_precompute(u64p(out.reshape(-1)), wbits, u64p(pts.reshape(-1)), n)
    return out


def p2_mult_wbits(table, wbits, scalars, n, nbits):
    r = np.zeros(36, np.uint64)
    tab = np.ascontiguousarray(table, dtype=np.uint64).reshape(-1)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
    lib().or_p2_mult_wbits(u64p(r), u64p(tab), wbits, n, u8p(sc), nbits)
    return r


def table_booth(scalar, nbits, wbits, rem):
    s = np.ascontiguousarray(scalar, dtype=np.uint8)
    return int(lib().or_table_booth(u8p(s), nbits, wbits, rem))


def gen_p2_points(n, seed):
    out = np.zeros((n, 24), np.uint64)
    if n:
        lib().or_gen_p2_points(u64p(out.reshape(-1)), n, seed)
    return out


# ---- synthetic inputs ------------------------------------------------------

def gen_scalars(n, seed, low64_pct=0):
    out = np.zeros((n, 32), np.uint8)
    lib().or_gen_scalars(u8p(out.reshape(-1)), n, seed, low64_pct)
    return out


def gen_points(n, seed, threads=8):
    out = np.zeros((n, 12), np.uint64)
    lib().or_gen_points(u64p(out.reshape(-1)), n, seed, threads)
    return out


def gen_srs(n, seed, threads=8):
    out = np.zeros((n, 12), np.uint64)
    lib().or_gen_srs(u64p(out.reshape(-1)), n, seed, threads)
    return out


def apply_edge_cases(points, scalars, seed):
    n = points.shape[0]
    assert points.flags.c_contiguous and scalars.flags.c_contiguous and scalars.shape[1] == 32
    lib().or_apply_edge_cases(u64p(points.reshape(-1)), u8p(scalars.reshape(-1)), n, seed)
