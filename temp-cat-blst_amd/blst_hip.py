"""ctypes binding of libblst_hip.so, the gfx950 G1 MSM backend.

The product is the C ABI in include/blst_hip.h; this module only forwards
numpy buffers to it for tests and benchmarks.  It never falls back to a CPU
implementation: if the library or a GPU is missing, calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libblst_hip.so")
# test and workload-generation hooks (include/blst_hip_testkit.h)
TESTKIT_PATH = os.path.join(HERE, "libblst_hip_testkit.so")

# symbol -> (argtypes, restype); the full exported surface of include/blst_hip.h
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
SIGNATURES = {
    "blst_p1s_mult_pippenger_scratch_sizeof": ([_sz], _sz),
    "blst_p1s_mult_pippenger_window_scratch_sizeof": ([_sz], _sz),
    "blst_p1s_mult_pippenger": ([_vp, _vp, _sz, _vp, _sz, _vp], None),
    "blst_p1s_mult_pippenger_window": ([_vp, _vp, _sz, _vp, _sz, _vp, _sz], None),
    "blst_p1s_tile_pippenger": ([_vp, _vp, _sz, _vp, _sz, _vp, _sz, _sz], None),
    "blst_p1s_mult_pippenger_fr_scratch_sizeof": ([_sz], _sz),
    "blst_p1s_mult_pippenger_fr": ([_vp, _vp, _vp, _sz, _vp, _sz], ctypes.c_bool),
    "blst_p1s_mult_fr_scratch_sizeof": ([_sz], _sz),
    "blst_p1s_mult_fr": ([_vp, _vp, _vp, _sz, _vp, _sz], ctypes.c_bool),
    "blst_p1s_to_affine": ([_vp, _vp, _sz], None),
    "blst_p1s_mult_wbits_precompute_sizeof": ([_sz, _sz], _sz),
    "blst_p1s_mult_wbits_precompute": ([_vp, _sz, _vp, _sz], None),
    "blst_p1s_mult_wbits_scratch_sizeof": ([_sz], _sz),
    "blst_p1s_mult_wbits": ([_vp, _vp, _sz, _sz, _vp, _sz, _vp], None),
    "blst_p2s_to_affine": ([_vp, _vp, _sz], None),
    "blst_p2s_mult_wbits_precompute_sizeof": ([_sz, _sz], _sz),
    "blst_p2s_mult_wbits_precompute": ([_vp, _sz, _vp, _sz], None),
    "blst_p2s_mult_wbits_scratch_sizeof": ([_sz], _sz),
    "blst_p2s_mult_wbits": ([_vp, _vp, _sz, _sz, _vp, _sz, _vp], None),
    "blst_hip_p1s_mult_wbits_precompute_device": ([_vp, _sz, _vp, _sz], ctypes.c_int),
    "blst_hip_p1s_mult_wbits_device": ([_vp, _vp, _sz, _sz, _vp, _sz], ctypes.c_int),
    "blst_hip_comb_table_bytes": ([_sz, _sz, _sz], _sz),
    "blst_hip_comb_table_build": ([_vp, _sz, _sz, _vp, _sz], ctypes.c_int),
    "blst_hip_comb_msm_batch": ([_vp, _vp, _sz, _sz, _vp, _sz, _sz], ctypes.c_int),
    "blst_hip_p1s_to_affine_device": ([_vp, _vp, _sz], ctypes.c_int),
    "blst_p2s_mult_pippenger_scratch_sizeof": ([_sz], _sz),
    "blst_p2s_mult_pippenger_window_scratch_sizeof": ([_sz], _sz),
    "blst_p2s_mult_pippenger": ([_vp, _vp, _sz, _vp, _sz, _vp], None),
    "blst_p2s_mult_pippenger_window": ([_vp, _vp, _sz, _vp, _sz, _vp, _sz], None),
    "blst_p2s_tile_pippenger": ([_vp, _vp, _sz, _vp, _sz, _vp, _sz, _sz], None),
    "blst_hip_p2_msm_device": ([_vp, _vp, _vp, _sz, _sz, _sz], ctypes.c_int),
    "blst_p2s_mult_pippenger_fr_scratch_sizeof": ([_sz], _sz),
    "blst_p2s_mult_pippenger_fr": ([_vp, _vp, _vp, _sz, _vp, _sz], ctypes.c_bool),
    "blst_p2s_mult_fr_scratch_sizeof": ([_sz], _sz),
    "blst_p2s_mult_fr": ([_vp, _vp, _vp, _sz, _vp, _sz], ctypes.c_bool),
    "blst_hip_last_error": ([], ctypes.c_char_p),
    "blst_hip_available": ([], ctypes.c_int),
    "blst_hip_msm_device": ([_vp, _vp, _vp, _sz, _sz, _sz], ctypes.c_int),
    "blst_hip_msm_device_noalloc": ([_vp, _vp, _vp, _sz, _sz], ctypes.c_int),
    "blst_hip_set_pipeline": ([_sz], ctypes.c_int),
    "blst_hip_set_pipeline_share": ([_sz], ctypes.c_int),
    "blst_hip_last_phase_ms": ([_vp], None),
    "blst_hip_device_upload": ([_vp, _sz], _vp),
    "blst_hip_device_free": ([_vp], None),
    "blst_hip_device_write": ([_vp, _vp, _sz], ctypes.c_int),
    "blst_hip_device_read": ([_vp, _vp, _sz], ctypes.c_int),
    "blst_hip_device_alloc": ([_sz], _vp),
    "blst_hip_p1_compress": ([_vp, _vp], None),
    "blst_hip_msm_batch_device": ([_vp, _vp, _sz], ctypes.c_int),
    "blst_hip_batch_plan_create": ([_vp, _sz], _vp),
    "blst_hip_batch_plan_run": ([_vp, _vp], ctypes.c_int),
    "blst_hip_batch_plan_destroy": ([_vp], None),
    "blst_hip_server_create": ([_sz, _sz], _vp),
    "blst_hip_server_destroy": ([_vp], None),
    "blst_hip_server_add_bases": ([_vp, _vp, _sz], ctypes.c_int),
    "blst_hip_server_submit": ([_vp, ctypes.c_int, _vp, _vp, _sz, _sz], ctypes.c_int64),
    "blst_hip_server_wait": ([_vp, ctypes.c_int64, _vp, _vp], ctypes.c_int),
    "blst_hip_server_stats": ([_vp, _vp], None),
    "blst_hip_host_alloc": ([_sz], _vp),
    "blst_hip_host_free": ([_vp], None),
    "blst_hip_msm_device_tuned": ([_vp, _vp, _vp, _sz, _sz, _sz, _sz, _sz], ctypes.c_int),
    "blst_hip_tuning_set": ([ctypes.c_int, _sz, _sz, _sz, _sz], ctypes.c_int),
    "blst_hip_tuning_get": ([ctypes.c_int, _sz, _sz, _vp], None),
}

TESTKIT_SIGNATURES = {
    "blst_hip_fp_ops": ([ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _sz], ctypes.c_int),
    "blst_hip_fq_ops": ([ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _sz], ctypes.c_int),
    "blst_hip_set_alloc_cap": ([_sz], None),
    "blst_hip_set_stage_failure": ([ctypes.c_int], None),
    "blst_hip_mad_issue_rate": ([_vp], ctypes.c_int),
    "blst_hip_batch_affine_probe": ([_vp, _vp, _sz, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
    "blst_hip_gather_probe": ([_vp, _sz, _vp, _sz, _vp, _vp], ctypes.c_int),
    "blst_hip_booth_digits": ([_vp, _vp, _sz, _sz, _sz, _sz], ctypes.c_int),
    "blst_hip_point_ops_check": ([_vp, _sz], ctypes.c_int),
    "blst_hip_p1s_mult_generator": ([_vp, _vp, _sz, _sz], ctypes.c_int),
    "blst_hip_synth_scalars": ([_vp, _sz, ctypes.c_uint64, ctypes.c_int], None),
    "blst_hip_synth_srs_scalars": ([_vp, _sz, ctypes.c_uint64], None),
    "blst_hip_synth_edge_cases": ([_vp, _vp, _sz, ctypes.c_uint64], None),
    "blst_hip_synth_fr": ([_vp, _vp, _sz], None),
    "blst_hip_synth_jacobian": ([_vp, _vp, _sz, ctypes.c_uint64], None),
    "blst_hip_synth_p2_points": ([_vp, _sz, ctypes.c_uint64], None),
}

_lib = None
_kit = None


def _bind(path, sigs, what):
    if not os.path.exists(path):
        raise RuntimeError("%s is not built (run __graft_entry__.build())" % what)
    L = ctypes.CDLL(path)
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


def load():
    """Load the library (no GPU call is made here)."""
    global _lib
    if _lib is None:
        _lib = _bind(LIB_PATH, SIGNATURES, "libblst_hip.so")
    return _lib


def testkit():
    """The test and workload-generation hooks (loads the product first)."""
    global _kit
    if _kit is None:
        load()
        _kit = _bind(TESTKIT_PATH, TESTKIT_SIGNATURES, "libblst_hip_testkit.so")
    return _kit


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def last_error():
    return load().blst_hip_last_error().decode()


def available():
    return bool(load().blst_hip_available())


def scratch_sizeof(n):
    return load().blst_p1s_mult_pippenger_scratch_sizeof(n)


def window_scratch_sizeof(w):
    return load().blst_p1s_mult_pippenger_window_scratch_sizeof(w)


class Descriptors:
    """Pointer arrays in the reference's descriptor convention.

    mode "contiguous": {base, NULL} (the Rust binding's form);
    mode "exact": n explicit pointers, no sentinel;
    mode "mixed": explicit pointers for the first k entries, then NULL.
    """

    def __init__(self, points, scalars, mode="contiguous", k=1):
        self.points = np.ascontiguousarray(points, dtype=np.uint64)
        self.scalars = np.ascontiguousarray(scalars, dtype=np.uint8)
        n = self.points.shape[0]
        nb = self.scalars.shape[1]
        pbase, sbase = self.points.ctypes.data, self.scalars.ctypes.data
        if mode == "contiguous":
            # entries past the first NULL are never read
            pp, ss = [pbase, 0], [sbase, 0]
        elif mode == "exact":
            pp = [pbase + 96 * i for i in range(n)]
            ss = [sbase + nb * i for i in range(n)]
        elif mode == "mixed":
            pp = [pbase + 96 * i if i < k else 0 for i in range(n)]
            ss = [sbase + nb * i if i < k else 0 for i in range(n)]
        else:
            raise ValueError(mode)
        self.pp = np.array(pp, dtype=np.uint64)
        self.ss = np.array(ss, dtype=np.uint64)


def p1s_mult_pippenger(points, scalars, nbits, window=0, scratch=True, mode="contiguous", k=1):
    """G1 MSM through the drop-in entry point; returns the 18-limb blst_p1."""
    L = load()
    n = points.shape[0]
    d = Descriptors(points, scalars, mode, k)
    ret = np.zeros(18, np.uint64)
    sz = (window_scratch_sizeof(window) if window else scratch_sizeof(n)) or 8
    scr = np.zeros((sz + 7) // 8, np.uint64) if scratch else None
    sp = _ptr(scr) if scratch else None
    if window:
        L.blst_p1s_mult_pippenger_window(_ptr(ret), _ptr(d.pp), n, _ptr(d.ss), nbits, sp, window)
    else:
        L.blst_p1s_mult_pippenger(_ptr(ret), _ptr(d.pp), n, _ptr(d.ss), nbits, sp)
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p1s_mult_pippenger: " + err)
    return ret


def fr_scratch_sizeof(n, entry="pippenger"):
    """Scratch bytes of blst_p1s_mult_pippenger_fr (entry "pippenger") or of
    blst_p1s_mult_fr (entry "msm")."""
    L = load()
    if entry == "pippenger":
        return L.blst_p1s_mult_pippenger_fr_scratch_sizeof(n)
    if entry == "msm":
        return L.blst_p1s_mult_fr_scratch_sizeof(n)
    raise ValueError(entry)


def p1s_mult_fr(points, fr, entry="pippenger", scratch_size=None, ret=None, null=()):
    """G1 MSM of contiguous points and Montgomery-form Fr scalars (n x 4
    uint64 limbs) through blst_p1s_mult_pippenger_fr / blst_p1s_mult_fr.

    scratch_size defaults to the queried size; null names arguments passed as
    NULL ("ret", "points", "scalars", "scratch").  Returns (ok, ret): ret is
    the 18-limb blst_p1 buffer, untouched by a call that returns False."""
    L = load()
    points = np.ascontiguousarray(points, dtype=np.uint64)
    fr = np.ascontiguousarray(fr, dtype=np.uint64)
    n = points.shape[0]
    assert fr.shape[0] == n
    if ret is None:
        ret = np.zeros(18, np.uint64)
    need = fr_scratch_sizeof(n, entry)
    if scratch_size is None:
        scratch_size = need
    scr = np.zeros(max((scratch_size + 7) // 8, 1), np.uint64)
    args = {"ret": _ptr(ret), "points": _ptr(points) if n else _ptr(np.zeros(12, np.uint64)),
            "scalars": _ptr(fr) if n else _ptr(np.zeros(4, np.uint64)), "scratch": _ptr(scr)}
    for k in null:
        args[k] = None
    f = L.blst_p1s_mult_pippenger_fr if entry == "pippenger" else L.blst_p1s_mult_fr
    ok = f(args["ret"], args["points"], args["scalars"], n, args["scratch"], scratch_size)
    return bool(ok), ret


def p2_fr_scratch_sizeof(n, entry="pippenger"):
    L = load()
    f = L.blst_p2s_mult_pippenger_fr_scratch_sizeof if entry == "pippenger" else L.blst_p2s_mult_fr_scratch_sizeof
    return f(n)


def p2s_mult_fr(points, fr, entry="pippenger", scratch_size=None, ret=None, null=()):
    """G2 MSM of contiguous points (n x 24 uint64) and Montgomery-form Fr
    scalars through blst_p2s_mult_pippenger_fr / blst_p2s_mult_fr; returns
    (ok, 36-limb blst_p2 buffer), the buffer untouched by a call that returns
    False."""
    L = load()
    points = np.ascontiguousarray(points, dtype=np.uint64)
    fr = np.ascontiguousarray(fr, dtype=np.uint64)
    n = points.shape[0]
    assert fr.shape[0] == n
    if ret is None:
        ret = np.zeros(36, np.uint64)
    need = p2_fr_scratch_sizeof(n, entry)
    if scratch_size is None:
        scratch_size = need
    scr = np.zeros(max((scratch_size + 7) // 8, 1), np.uint64)
    args = {"ret": _ptr(ret), "points": _ptr(points) if n else _ptr(np.zeros(24, np.uint64)),
            "scalars": _ptr(fr) if n else _ptr(np.zeros(4, np.uint64)), "scratch": _ptr(scr)}
    for k in null:
        args[k] = None
    f = L.blst_p2s_mult_pippenger_fr if entry == "pippenger" else L.blst_p2s_mult_fr
    ok = f(args["ret"], args["points"], args["scalars"], n, args["scratch"], scratch_size)
    return bool(ok), ret


def p1s_to_affine(points, mode="contiguous", k=1):
    """(n, 18) uint64 Jacobian points -> (n, 12) affine through
    blst_p1s_to_affine, descriptors as in Descriptors."""
    L = load()
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    n = pts.shape[0]
    base = pts.ctypes.data
    if mode == "contiguous":
        pp = [base, 0]
    elif mode == "exact":
        pp = [base + 144 * i for i in range(n)]
    elif mode == "mixed":
        pp = [base + 144 * i if i < k else 0 for i in range(n)]
    else:
        raise ValueError(mode)
    pp = np.array(pp, dtype=np.uint64)
    out = np.zeros((max(n, 1), 12), np.uint64)
    L.blst_p1s_to_affine(_ptr(out), _ptr(pp), n)
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p1s_to_affine: " + err)
    return out[:n]


def p1s_to_affine_device(d_out_ptr, d_points_ptr, n):
    if load().blst_hip_p1s_to_affine_device(d_out_ptr, d_points_ptr, n) != 0:
        raise RuntimeError("blst_hip_p1s_to_affine_device: " + last_error())


def wbits_precompute(points, wbits, mode="contiguous", k=1):
    """(n, 12) affine points -> (n * 2^(wbits-1), 12) table through
    blst_p1s_mult_wbits_precompute (descriptors as in Descriptors)."""
    L = load()
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    n = pts.shape[0]
    d = Descriptors(pts, np.zeros((n, 1), np.uint8), mode, k)
    size = L.blst_p1s_mult_wbits_precompute_sizeof(wbits, n)
    table = np.zeros((max(size // 96, 1), 12), np.uint64)
    L.blst_p1s_mult_wbits_precompute(_ptr(table), wbits, _ptr(d.pp), n)
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p1s_mult_wbits_precompute: " + err)
    return table[:size // 96]


def mult_wbits(table, wbits, scalars, nbits, mode="contiguous", k=1):
    """MSM over a wbits table through blst_p1s_mult_wbits."""
    L = load()
    sc = np.ascontiguousarray(scalars, dtype=np.uint8)
    n = sc.shape[0]
    tab = np.ascontiguousarray(table, dtype=np.uint64)
    d = Descriptors(np.zeros((n, 12), np.uint64), sc, mode, k)
    scr = np.zeros(max(L.blst_p1s_mult_wbits_scratch_sizeof(n) // 8, 1), np.uint64)
    ret = np.zeros(18, np.uint64)
    L.blst_p1s_mult_wbits(_ptr(ret), _ptr(tab), wbits, n, _ptr(d.ss), nbits, _ptr(scr))
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p1s_mult_wbits: " + err)
    return ret


def _descriptor_array(base, n, stride, mode, k):
    if mode == "contiguous":
        return np.array([base, 0], np.uint64)
    if mode == "exact":
        return np.array([base + stride * i for i in range(n)], np.uint64)
    if mode == "mixed":
        return np.array([base + stride * i if i < k else 0 for i in range(n)], np.uint64)
    raise ValueError(mode)


def p2s_to_affine(points, mode="contiguous", k=1):
    """(n, 36) uint64 G2 Jacobian points -> (n, 24) affine through
    blst_p2s_to_affine."""
    L = load()
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    n = pts.shape[0]
    pp = _descriptor_array(pts.ctypes.data, n, 288, mode, k)
    out = np.zeros((max(n, 1), 24), np.uint64)
    L.blst_p2s_to_affine(_ptr(out), _ptr(pp), n)
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p2s_to_affine: " + err)
    return out[:n]


def p2_wbits_precompute(points, wbits, mode="contiguous", k=1):
    """(n, 24) G2 affine points -> (n * 2^(wbits-1), 24) table through
    blst_p2s_mult_wbits_precompute."""
    L = load()
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    n = pts.shape[0]
    pp = _descriptor_array(pts.ctypes.data, n, 192, mode, k)
    size = L.blst_p2s_mult_wbits_precompute_sizeof(wbits, n)
    table = np.zeros((max(size // 192, 1), 24), np.uint64)
    L.blst_p2s_mult_wbits_precompute(_ptr(table), wbits, _ptr(pp), n)
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p2s_mult_wbits_precompute: " + err)
    return table[:size // 192]


def p2_mult_wbits(table, wbits, scalars, nbits, mode="contiguous", k=1):
    """G2 MSM over a wbits table through blst_p2s_mult_wbits."""
    L = load()
    sc = np.ascontiguousarray(scalars, dtype=np.uint8)
    n = sc.shape[0]
    tab = np.ascontiguousarray(table, dtype=np.uint64)
    ss = _descriptor_array(sc.ctypes.data, n, sc.shape[1], mode, k)
    scr = np.zeros(max(L.blst_p2s_mult_wbits_scratch_sizeof(n) // 8, 1), np.uint64)
    ret = np.zeros(36, np.uint64)
    L.blst_p2s_mult_wbits(_ptr(ret), _ptr(tab), wbits, n, _ptr(ss), nbits, _ptr(scr))
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p2s_mult_wbits: " + err)
    return ret


def wbits_precompute_device(d_table_ptr, wbits, d_points_ptr, n):
    if load().blst_hip_p1s_mult_wbits_precompute_device(d_table_ptr, wbits, d_points_ptr, n) != 0:
        raise RuntimeError("blst_hip_p1s_mult_wbits_precompute_device: " + last_error())


def mult_wbits_device(d_table_ptr, wbits, n, d_scalars_ptr, nbits):
    ret = np.zeros(18, np.uint64)
    if load().blst_hip_p1s_mult_wbits_device(_ptr(ret), d_table_ptr, wbits, n, d_scalars_ptr, nbits) != 0:
        raise RuntimeError("blst_hip_p1s_mult_wbits_device: " + last_error())
    return ret


def comb_table_bytes(n, wbits, nbits=255):
    return load().blst_hip_comb_table_bytes(n, wbits, nbits)


def comb_table_build(d_table_ptr, wbits, d_points_ptr, n, nbits=255):
    if load().blst_hip_comb_table_build(d_table_ptr, wbits, nbits, d_points_ptr, n) != 0:
        raise RuntimeError("blst_hip_comb_table_build: " + last_error())


def comb_msm_batch(d_table_ptr, wbits, n, d_scalars_ptr, njobs, nbits=255):
    rets = np.zeros((njobs, 18), np.uint64)
    if load().blst_hip_comb_msm_batch(_ptr(rets), d_table_ptr, wbits, n, d_scalars_ptr, njobs, nbits) != 0:
        raise RuntimeError("blst_hip_comb_msm_batch: " + last_error())
    return rets


class DeviceAlloc:
    """Uninitialised device memory of a given byte size (freed on close)."""

    def __init__(self, nbytes):
        self.ptr = load().blst_hip_device_alloc(nbytes)
        if not self.ptr:
            raise RuntimeError("blst_hip_device_alloc: " + last_error())
        self.nbytes = nbytes

    def read(self):
        out = np.zeros(self.nbytes, np.uint8)
        if load().blst_hip_device_read(_ptr(out), ctypes.c_void_p(self.ptr), self.nbytes):
            raise RuntimeError("blst_hip_device_read: " + last_error())
        return out

    def close(self):
        if self.ptr:
            load().blst_hip_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = None


def p2s_mult_pippenger(points, scalars, nbits, window=0, scratch=True, mode="contiguous", k=1, bit0=None):
    """G2 MSM through the drop-in entry points (points: n x 24 uint64);
    bit0 selects blst_p2s_tile_pippenger.  Returns the 36-limb blst_p2."""
    L = load()
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    sc = np.ascontiguousarray(scalars, dtype=np.uint8)
    n = pts.shape[0]
    nb = sc.shape[1] if sc.ndim == 2 else 1
    base_p, base_s = pts.ctypes.data, sc.ctypes.data
    if mode == "contiguous":
        pp, ss = [base_p, 0], [base_s, 0]
    elif mode == "exact":
        pp = [base_p + 192 * i for i in range(n)]
        ss = [base_s + nb * i for i in range(n)]
    elif mode == "mixed":
        pp = [base_p + 192 * i if i < k else 0 for i in range(n)]
        ss = [base_s + nb * i if i < k else 0 for i in range(n)]
    else:
        raise ValueError(mode)
    pp = np.array(pp, np.uint64)
    ss = np.array(ss, np.uint64)
    ret = np.zeros(36, np.uint64)
    sz = (L.blst_p2s_mult_pippenger_window_scratch_sizeof(window) if window
          else L.blst_p2s_mult_pippenger_scratch_sizeof(n)) or 8
    scr = np.zeros((sz + 7) // 8, np.uint64) if scratch else None
    sp = _ptr(scr) if scratch else None
    if bit0 is not None:
        L.blst_p2s_tile_pippenger(_ptr(ret), _ptr(pp), n, _ptr(ss), nbits, sp, bit0, window)
    elif window:
        L.blst_p2s_mult_pippenger_window(_ptr(ret), _ptr(pp), n, _ptr(ss), nbits, sp, window)
    else:
        L.blst_p2s_mult_pippenger(_ptr(ret), _ptr(pp), n, _ptr(ss), nbits, sp)
    err = last_error()
    if err and n > 0:
        raise RuntimeError("blst_p2s_mult_pippenger: " + err)
    return ret


def p2_msm_device(d_points_ptr, d_scalars_ptr, n, nbits, window=0):
    ret = np.zeros(36, np.uint64)
    if load().blst_hip_p2_msm_device(_ptr(ret), d_points_ptr, d_scalars_ptr, n, nbits, window) != 0:
        raise RuntimeError("blst_hip_p2_msm_device: " + last_error())
    return ret


def p1s_tile_pippenger(points, scalars, nbits, bit0, window):
    L = load()
    n = points.shape[0]
    d = Descriptors(points, scalars, "contiguous")
    ret = np.zeros(18, np.uint64)
    scr = np.zeros((window_scratch_sizeof(window) + 7) // 8 + 1, np.uint64)
    L.blst_p1s_tile_pippenger(_ptr(ret), _ptr(d.pp), n, _ptr(d.ss), nbits, _ptr(scr), bit0, window)
    err = last_error()
    if err:
        raise RuntimeError("blst_p1s_tile_pippenger: " + err)
    return ret


FP_OPS = {"mul": 0, "sqr": 1, "mul_sum": 2, "add": 3, "sub": 4, "dbl": 5, "neg": 6, "cneg": 7, "canon": 8,
          "is_zero": 9, "inv": 10}


def fp_ops(op, a, b=None, c=None, d=None):
    """One device field operation over arrays of (n, 12) uint32 limbs (or
    (n, 6) uint64): returns (n, 12) uint32."""
    A = np.ascontiguousarray(a).view(np.uint32).reshape(-1, 12)
    n = A.shape[0]
    B = np.ascontiguousarray(b).view(np.uint32).reshape(-1, 12) if b is not None else A
    C = np.ascontiguousarray(c).view(np.uint32).reshape(-1, 12) if c is not None else A
    D = np.ascontiguousarray(d).view(np.uint32).reshape(-1, 12) if d is not None else A
    for X in (B, C, D):
        assert X.shape[0] == n
    out = np.zeros((n, 12), np.uint32)
    rc = testkit().blst_hip_fp_ops(FP_OPS[op], _ptr(out), _ptr(A), _ptr(B), _ptr(C), _ptr(D), n)
    if rc:
        raise RuntimeError("blst_hip_fp_ops: " + last_error())
    return out


FQ_OPS = {"mul": 0, "sqr": 1, "mul_sum": 2, "reduce": 3, "sub_lazy": 4, "neg_lazy": 5, "x3": 6, "is_zero": 7,
          "from_words8": 8, "from_words4": 9, "from_words0": 10, "to_words": 11, "times4": 12,
          "reduce_wide": 13, "canon": 14}


def fq_ops(op, a, b=None, c=None, d=None):
    """One device operation of the G1 accumulation's 28-bit limb field over
    arrays of (n, 14) uint32 limbs: returns (n, 14) uint32."""
    A = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 14)
    n = A.shape[0]
    B, C, D = (np.ascontiguousarray(x, dtype=np.uint32).reshape(-1, 14) if x is not None else A for x in (b, c, d))
    for X in (B, C, D):
        assert X.shape[0] == n
    out = np.zeros((n, 14), np.uint32)
    rc = testkit().blst_hip_fq_ops(FQ_OPS[op], _ptr(out), _ptr(A), _ptr(B), _ptr(C), _ptr(D), n)
    if rc:
        raise RuntimeError("blst_hip_fq_ops: " + last_error())
    return out


def gather_probe(d_points_ptr, stride_bytes, d_index_ptr, nentries, read=True):
    """The accumulation's gathers at a 96- or 128-byte point stride: returns
    (ns per entry, (nentries // 32, 24) uint64 canonical XYZZ sums or None)."""
    d_out = DeviceAlloc((nentries // 32) * 192)
    try:
        ns = np.zeros(1, np.float64)
        rc = testkit().blst_hip_gather_probe(ctypes.c_void_p(d_points_ptr), stride_bytes, ctypes.c_void_p(d_index_ptr),
                                             nentries, _ptr(ns), ctypes.c_void_p(d_out.ptr))
        if rc:
            raise RuntimeError("blst_hip_gather_probe: " + last_error())
        return float(ns[0]), (d_out.read().view(np.uint64).reshape(-1, 24) if read else None)
    finally:
        d_out.close()


def mad_issue_rate():
    """Measured v_mad_u64_u32 multiply-adds per second of the device."""
    out = ctypes.c_double(0)
    if testkit().blst_hip_mad_issue_rate(ctypes.byref(out)):
        raise RuntimeError("blst_hip_mad_issue_rate: " + last_error())
    return out.value


def batch_affine_probe(d_points_ptr, npairs, pairs_per_lane, d_index_ptr=None, read=True):
    """Times batch-affine and XYZZ additions of the same pairs; returns
    (ns per pair [batch-affine, xyzz], affine sums (npairs, 12) uint64,
    xyzz sums (npairs, 24) uint64), the sums None when read is False."""
    d_aff = DeviceAlloc(npairs * 96)
    d_xz = DeviceAlloc(npairs * 192)
    try:
        ns = np.zeros(2, np.float64)
        rc = testkit().blst_hip_batch_affine_probe(ctypes.c_void_p(d_points_ptr),
                                                   ctypes.c_void_p(d_index_ptr) if d_index_ptr else None, npairs,
                                                   pairs_per_lane, _ptr(ns), ctypes.c_void_p(d_aff.ptr),
                                                   ctypes.c_void_p(d_xz.ptr))
        if rc:
            raise RuntimeError("blst_hip_batch_affine_probe: " + last_error())
        aff = d_aff.read().view(np.uint64).reshape(npairs, 12) if read else None
        xz = d_xz.read().view(np.uint64).reshape(npairs, 24) if read else None
    finally:
        d_aff.close()
        d_xz.close()
    return ns, aff, xz


def set_alloc_cap(nbytes):
    """Device allocations of the product above nbytes fail (0: no cap)."""
    testkit().blst_hip_set_alloc_cap(nbytes)


def set_stage_failure(range_index):
    """The staging thread of split host-buffer MSMs fails before this point
    range (-1: never)."""
    testkit().blst_hip_set_stage_failure(range_index)


def booth_digits(scalars, nbits, bit0, window):
    L = testkit()
    sc = np.ascontiguousarray(scalars, dtype=np.uint8)
    out = np.zeros(sc.shape[0], np.int32)
    rc = L.blst_hip_booth_digits(_ptr(out), _ptr(sc), sc.shape[0], nbits, bit0, window)
    if rc:
        raise RuntimeError("blst_hip_booth_digits: " + last_error())
    return out


def set_pipeline(groups):
    """Window-group pipelines per lone MSM (0 = automatic)."""
    if load().blst_hip_set_pipeline(groups):
        raise RuntimeError("blst_hip_set_pipeline: " + last_error())


def set_pipeline_share(percent):
    """Percent of the windows in the last window group (0: equal groups)."""
    if load().blst_hip_set_pipeline_share(percent):
        raise RuntimeError("blst_hip_set_pipeline_share: " + last_error())


def point_ops_check(d_points_ptr, n):
    """Mismatching coordinates between the four-lane and one-lane point
    formulas on n device-resident affine points."""
    rc = testkit().blst_hip_point_ops_check(d_points_ptr, n)
    if rc < 0:
        raise RuntimeError("blst_hip_point_ops_check: " + last_error())
    return rc


def msm_device(d_points_ptr, d_scalars_ptr, n, nbits, window=0, chunk=0, segment=0):
    """MSM over device-resident inputs given as raw device addresses."""
    L = load()
    ret = np.zeros(18, np.uint64)
    rc = L.blst_hip_msm_device_tuned(_ptr(ret), ctypes.c_void_p(d_points_ptr), ctypes.c_void_p(d_scalars_ptr),
                                     n, nbits, window, chunk, segment)
    if rc:
        raise RuntimeError("blst_hip_msm_device: " + last_error())
    return ret


def compress(p1):
    """48-byte compressed encoding of an 18-limb blst_p1 (library code)."""
    p = np.ascontiguousarray(p1, dtype=np.uint64)
    out = np.zeros(48, np.uint8)
    load().blst_hip_p1_compress(_ptr(out), _ptr(p))
    return out.tobytes()


def generator_multiples(scalars, nbits=255):
    """[k_i]G on the device, returned as (n, 12) uint64 affine limbs."""
    sc = np.ascontiguousarray(scalars, dtype=np.uint8)
    out = np.zeros((sc.shape[0], 12), np.uint64)
    rc = testkit().blst_hip_p1s_mult_generator(_ptr(out), _ptr(sc), sc.shape[0], nbits)
    if rc:
        raise RuntimeError("blst_hip_p1s_mult_generator: " + last_error())
    return out


def synth_scalars(n, seed, low64_pct=0):
    out = np.zeros((n, 32), np.uint8)
    testkit().blst_hip_synth_scalars(_ptr(out), n, seed, low64_pct)
    return out


def synth_srs_scalars(n, seed):
    out = np.zeros((n, 32), np.uint8)
    testkit().blst_hip_synth_srs_scalars(_ptr(out), n, seed)
    return out


def synth_edge_cases(points, scalars, seed):
    assert points.flags.c_contiguous and scalars.flags.c_contiguous
    testkit().blst_hip_synth_edge_cases(_ptr(points), _ptr(scalars), points.shape[0], seed)


def synth_fr(scalars):
    """Montgomery-form blst_fr (n x 4 uint64) of canonical scalars below r."""
    sc = np.ascontiguousarray(scalars, dtype=np.uint8)
    out = np.zeros((sc.shape[0], 4), np.uint64)
    testkit().blst_hip_synth_fr(_ptr(out), _ptr(sc), sc.shape[0])
    return out


def synth_jacobian(affine, seed):
    """Jacobian representatives (n x 18 uint64) of affine points at seeded Z."""
    a = np.ascontiguousarray(affine, dtype=np.uint64)
    out = np.zeros((a.shape[0], 18), np.uint64)
    testkit().blst_hip_synth_jacobian(_ptr(out), _ptr(a), a.shape[0], seed)
    return out


def synth_p2_points(n, seed):
    """n G2 bases (n x 24 uint64 affine) in arithmetic progression [a + i b]G2."""
    out = np.zeros((n, 24), np.uint64)
    testkit().blst_hip_synth_p2_points(_ptr(out), n, seed)
    return out


def synth_points(n, seed):
    """P_i = [k_i]G with k_i uniform in Fr from the seeded stream."""
    return generator_multiples(synth_scalars(n, seed))


class MsmJob(ctypes.Structure):
    """blst_hip_msm_job"""
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("npoints", ctypes.c_size_t),
                ("nbits", ctypes.c_size_t), ("window", ctypes.c_size_t)]


def msm_batch_device(jobs):
    """jobs: sequence of (device_points, device_scalars, n, nbits[, window]).
    Returns an (njobs, 18) uint64 array of blst_p1 results."""
    arr = (MsmJob * len(jobs))()
    for k, j in enumerate(jobs):
        arr[k].points, arr[k].scalars, arr[k].npoints, arr[k].nbits = j[0], j[1], j[2], j[3]
        arr[k].window = j[4] if len(j) > 4 else 0
    out = np.zeros((len(jobs), 18), np.uint64)
    rc = load().blst_hip_msm_batch_device(_ptr(out), ctypes.cast(arr, ctypes.c_void_p), len(jobs))
    if rc:
        raise RuntimeError("blst_hip_msm_batch_device: " + last_error())
    return out


def _job_array(jobs):
    arr = (MsmJob * max(len(jobs), 1))()
    for k, j in enumerate(jobs):
        arr[k].points, arr[k].scalars, arr[k].npoints, arr[k].nbits = j[0], j[1], j[2], j[3]
        arr[k].window = j[4] if len(j) > 4 else 0
    return arr


class BatchPlan:
    """A batch captured into a HIP graph (blst_hip_batch_plan_*)."""

    def __init__(self, jobs):
        self.njobs = len(jobs)
        self._arr = _job_array(jobs)
        self.ptr = load().blst_hip_batch_plan_create(ctypes.cast(self._arr, ctypes.c_void_p), self.njobs)
        if not self.ptr:
            raise RuntimeError("blst_hip_batch_plan_create: " + last_error())

    def run(self):
        out = np.zeros((self.njobs, 18), np.uint64)
        rc = load().blst_hip_batch_plan_run(ctypes.c_void_p(self.ptr), _ptr(out))
        if rc:
            raise RuntimeError("blst_hip_batch_plan_run: " + last_error())
        return out

    def close(self):
        if self.ptr:
            load().blst_hip_batch_plan_destroy(ctypes.c_void_p(self.ptr))
            self.ptr = None


class PinnedArray:
    """A numpy view of page-locked host memory (freed on close)."""

    def __init__(self, shape, dtype):
        self.nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        self.ptr = load().blst_hip_host_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise RuntimeError("blst_hip_host_alloc: " + last_error())
        buf = (ctypes.c_uint8 * max(self.nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=self.nbytes).view(dtype).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            load().blst_hip_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = None


class Server:
    """The continuous-batching serving loop (blst_hip_server_*)."""

    def __init__(self, hbm_budget=0, max_batch_points=0):
        self.ptr = load().blst_hip_server_create(hbm_budget, max_batch_points)
        if not self.ptr:
            raise RuntimeError("blst_hip_server_create: " + last_error())
        self._keep = {}

    def add_bases(self, points):
        points = np.ascontiguousarray(points, dtype=np.uint64)
        bid = load().blst_hip_server_add_bases(ctypes.c_void_p(self.ptr), _ptr(points), points.shape[0])
        if bid < 0:
            raise RuntimeError("blst_hip_server_add_bases: " + last_error())
        return bid

    def submit(self, scalars, nbits, base_id=-1, points=None, n=None):
        """scalars/points must stay alive until wait(); kept referenced here."""
        n = scalars.shape[0] if n is None else n
        pp = _ptr(points) if points is not None else None
        t = load().blst_hip_server_submit(ctypes.c_void_p(self.ptr), base_id, pp, _ptr(scalars), n, nbits)
        if t < 0:
            raise RuntimeError("blst_hip_server_submit: " + last_error())
        self._keep[t] = (scalars, points)
        return t

    def wait(self, ticket):
        ret = np.zeros(18, np.uint64)
        lat = ctypes.c_double(0)
        rc = load().blst_hip_server_wait(ctypes.c_void_p(self.ptr), ticket, _ptr(ret), ctypes.byref(lat))
        self._keep.pop(ticket, None)
        if rc:
            raise RuntimeError("blst_hip_server_wait: " + last_error())
        return ret, lat.value

    def stats(self):
        out = np.zeros(6, np.uint64)
        load().blst_hip_server_stats(ctypes.c_void_p(self.ptr), _ptr(out))
        return dict(zip(["submitted", "completed", "batches", "max_inflight", "peak_bytes", "budget"],
                        [int(x) for x in out]))

    def close(self):
        if self.ptr:
            load().blst_hip_server_destroy(ctypes.c_void_p(self.ptr))
            self.ptr = None


class DeviceBuffer:
    """A device copy of a host numpy array (freed on close)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.ptr = load().blst_hip_device_upload(_ptr(host), host.nbytes)
        if not self.ptr:
            raise RuntimeError("blst_hip_device_upload: " + last_error())

    def write(self, host):
        """Overwrite the buffer's contents (same byte size) from host memory."""
        host = np.ascontiguousarray(host)
        assert host.nbytes == self.nbytes
        if load().blst_hip_device_write(ctypes.c_void_p(self.ptr), _ptr(host), host.nbytes):
            raise RuntimeError("blst_hip_device_write: " + last_error())

    def read(self):
        """The buffer's contents as a uint8 host array."""
        out = np.zeros(self.nbytes, np.uint8)
        if load().blst_hip_device_read(_ptr(out), ctypes.c_void_p(self.ptr), self.nbytes):
            raise RuntimeError("blst_hip_device_read: " + last_error())
        return out

    def close(self):
        if self.ptr:
            load().blst_hip_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = None


def tuning_set(batch, log2n, window, chunk, segment):
    if load().blst_hip_tuning_set(int(batch), log2n, window, chunk, segment):
        raise RuntimeError("blst_hip_tuning_set: " + last_error())


def tuning_get(batch, n, nbits=255):
    out = np.zeros(3, np.uint64)
    load().blst_hip_tuning_get(int(batch), n, nbits, _ptr(out))
    return tuple(int(x) for x in out)


def last_phase_ms():
    out = np.zeros(7, np.float64)
    load().blst_hip_last_phase_ms(_ptr(out))
    return dict(zip(["digits", "sort", "accumulate", "partials", "reduce", "fold", "total"], out.tolist()))
