"""The C ABI boundary without a GPU: the library loads, exports every symbol
include/blst_hip.h declares, and its host-only entry points (scratch-size
queries, compression, workload generators) agree with the oracle."""
import os
import random
import re

import numpy as np

import pyref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="blst_hip.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(blst_\w+)\s*\(", text)))


def test_header_declares_drop_in_surface():
    syms = declared_symbols()
    for s in ("blst_p1s_mult_pippenger", "blst_p1s_mult_pippenger_scratch_sizeof",
              "blst_p1s_mult_pippenger_window", "blst_p1s_mult_pippenger_window_scratch_sizeof",
              "blst_p1s_tile_pippenger"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import ctypes
    import blst_hip
    L = ctypes.CDLL(blst_hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding binds exactly that surface
    assert sorted(blst_hip.SIGNATURES) == declared_symbols()


def test_testkit_exports_its_hooks_and_the_product_does_not():
    """The test and workload-generation hooks live in their own library: the
    product exports none of them."""
    import ctypes
    import blst_hip
    kit = declared_symbols("blst_hip_testkit.h")
    assert "blst_hip_fp_ops" in kit and "blst_hip_synth_scalars" in kit
    K = ctypes.CDLL(blst_hip.TESTKIT_PATH)
    assert not [s for s in kit if not hasattr(K, s)]
    assert sorted(blst_hip.TESTKIT_SIGNATURES) == kit
    assert not set(kit) & set(declared_symbols())
    L = ctypes.CDLL(blst_hip.LIB_PATH)
    assert not [s for s in kit if hasattr(L, s)]


def test_scratch_sizes_match_reference(orc):
    import blst_hip
    for n in [0, 1, 2, 3, 5, 16, 17, 100, 255, 256, 4095, 4096, 4097, 1 << 16, 1 << 18, 1 << 20, (1 << 24) + 3]:
        assert blst_hip.scratch_sizeof(n) == orc.scratch_sizeof(n), n
    for w in range(0, 30):
        assert blst_hip.window_scratch_sizeof(w) == orc.window_scratch_sizeof(w), w


def test_compress_matches_oracle(orc):
    import blst_hip
    rnd = random.Random(5)
    g = orc.from_affine(orc.generator_affine())
    for _ in range(20):
        k = rnd.getrandbits(255)
        p = orc.p1_mult(g, k.to_bytes(32, "little"), 255)
        assert blst_hip.compress(p) == orc.compress(p) == pyref.compress(pyref.mul(k, pyref.G))
    assert blst_hip.compress(np.zeros(18, np.uint64)) == bytes([0xC0]) + bytes(47)


def test_workload_generators_match_oracle(orc):
    import blst_hip
    assert np.array_equal(blst_hip.synth_scalars(300, 77), orc.gen_scalars(300, 77))
    assert np.array_equal(blst_hip.synth_scalars(300, 78, 10), orc.gen_scalars(300, 78, 10))
    s = blst_hip.synth_srs_scalars(5, 9)
    assert int.from_bytes(s[0].tobytes(), "little") == 1
    tau = int.from_bytes(s[1].tobytes(), "little")
    for i in range(5):
        assert int.from_bytes(s[i].tobytes(), "little") == pow(tau, i, pyref.R_ORDER)
    pts = orc.gen_points(500, 1, threads=2)
    sc = orc.gen_scalars(500, 2)
    p2, s2 = pts.copy(), sc.copy()
    orc.apply_edge_cases(pts, sc, 3)
    blst_hip.synth_edge_cases(p2, s2, 3)
    assert np.array_equal(pts, p2) and np.array_equal(sc, s2)


def test_fail_closed_without_gpu_work():
    """Argument checks that return before any device call (c_api.cpp:1390-1408)."""
    import blst_hip
    L = blst_hip.load()
    ret = np.full(18, 7, np.uint64)
    # npoints == 0 and nbits == 0 write infinity without reading inputs
    L.blst_p1s_mult_pippenger(blst_hip._ptr(ret), None, 0, None, 255, None)
    assert not ret.any()
    ret[:] = 7
    pts = np.zeros((2, 12), np.uint64)
    sc = np.zeros((2, 32), np.uint8)
    d = blst_hip.Descriptors(pts, sc)
    L.blst_p1s_mult_pippenger(blst_hip._ptr(ret), blst_hip._ptr(d.pp), 2, blst_hip._ptr(d.ss), 0, None)
    assert not ret.any()
    for nbits, scratch in ((257, True), (255, False)):
        ret[:] = 7
        scr = np.zeros(4096, np.uint64)
        L.blst_p1s_mult_pippenger(blst_hip._ptr(ret), blst_hip._ptr(d.pp), 2, blst_hip._ptr(d.ss), nbits,
                                  blst_hip._ptr(scr) if scratch else None)
        assert not ret.any()
    # null heads
    ret[:] = 7
    nullheads = np.zeros(2, np.uint64)
    scr = np.zeros(4096, np.uint64)
    L.blst_p1s_mult_pippenger(blst_hip._ptr(ret), blst_hip._ptr(nullheads), 2, blst_hip._ptr(d.ss), 255,
                              blst_hip._ptr(scr))
    assert not ret.any()
    # invalid windows for the window and tile variants
    for w in (0, 25):
        ret[:] = 7
        L.blst_p1s_mult_pippenger_window(blst_hip._ptr(ret), blst_hip._ptr(d.pp), 2, blst_hip._ptr(d.ss), 255,
                                         blst_hip._ptr(scr), w)
        assert not ret.any()
        ret[:] = 7
        L.blst_p1s_tile_pippenger(blst_hip._ptr(ret), blst_hip._ptr(d.pp), 2, blst_hip._ptr(d.ss), 255,
                                  blst_hip._ptr(scr), 0, w)
        assert not ret.any()
    # tile with bit0 >= nbits
    ret[:] = 7
    L.blst_p1s_tile_pippenger(blst_hip._ptr(ret), blst_hip._ptr(d.pp), 2, blst_hip._ptr(d.ss), 255,
                              blst_hip._ptr(scr), 255, 8)
    assert not ret.any()


def test_fr_scratch_sizes_match_oracle(orc):
    import blst_hip
    for n in [0, 1, 2, 3, 5, 64, 4095, 4096, 4097, 16384, 16385, 1 << 17, (1 << 18) - 1, 1 << 18,
              (1 << 18) + 1, 1 << 20, (1 << 32) + 1, 1 << 58]:
        assert blst_hip.fr_scratch_sizeof(n, "pippenger") == orc.pippenger_fr_scratch_sizeof(n), n
        assert blst_hip.fr_scratch_sizeof(n, "msm") == orc.msm_fr_scratch_sizeof(n), n
    # the full capacity range of the reference's monotonicity test
    L = blst_hip.load()
    prev = 0
    for n in range(1, (1 << 18) + 2):
        cur = L.blst_p1s_mult_fr_scratch_sizeof(n)
        assert cur >= prev, n
        prev = cur


def test_g2_fr_scratch_sizes_match_oracle(orc):
    import blst_hip
    for n in [0, 1, 2, 3, 5, 64, 4095, 4096, 4097, 16383, 16384, 16385, 1 << 17, 1 << 20, (1 << 32) + 1, 1 << 57]:
        assert blst_hip.p2_fr_scratch_sizeof(n, "pippenger") == orc.p2_pippenger_fr_scratch_sizeof(n), n
        assert blst_hip.p2_fr_scratch_sizeof(n, "msm") == orc.p2_msm_fr_scratch_sizeof(n), n


def test_g2_fr_fail_closed_without_gpu_work():
    """The G2 blst_fr entries keep the checks of c_api_pippenger_fr /
    c_api_msm_fr (c_api.cpp:532-552, 809-828): false with ret untouched, or
    true with infinity for no points."""
    import blst_hip
    pts = np.ones((4, 24), np.uint64)
    fr = np.ones((4, 4), np.uint64)
    for entry in ("pippenger", "msm"):
        for kw in ({"null": ("ret",)}, {"null": ("points",)}, {"null": ("scalars",)}, {"null": ("scratch",)},
                   {"scratch_size": blst_hip.p2_fr_scratch_sizeof(4, entry) - 1}, {"scratch_size": 0}):
            ret = np.full(36, 7, np.uint64)
            ok, out = blst_hip.p2s_mult_fr(pts, fr, entry, ret=ret, **kw)
            assert not ok and (out == 7).all(), (entry, kw)
        ret = np.full(36, 7, np.uint64)
        ok, out = blst_hip.p2s_mult_fr(pts[:0], fr[:0], entry, ret=ret, null=("points", "scalars", "scratch"))
        assert ok and not out.any()


def test_fr_fail_closed_without_gpu_work():
    """Argument checks of c_api_pippenger_fr / c_api_msm_fr
    (c_api.cpp:532-552, 809-828) that return before any device call: false
    with ret untouched, or true with infinity for no points."""
    import blst_hip
    pts = np.ones((4, 12), np.uint64)
    fr = np.ones((4, 4), np.uint64)
    for entry in ("pippenger", "msm"):
        for kw in ({"null": ("ret",)}, {"null": ("points",)}, {"null": ("scalars",)}, {"null": ("scratch",)},
                   {"scratch_size": blst_hip.fr_scratch_sizeof(4, entry) - 1}, {"scratch_size": 0}):
            ret = np.full(18, 7, np.uint64)
            ok, out = blst_hip.p1s_mult_fr(pts, fr, entry, ret=ret, **kw)
            assert not ok and (out == 7).all(), (entry, kw)
        ret = np.full(18, 7, np.uint64)
        ok, out = blst_hip.p1s_mult_fr(pts[:0], fr[:0], entry, ret=ret, null=("points", "scalars", "scratch"))
        assert ok and not out.any()


def test_to_affine_without_gpu_work():
    """blst_p1s_to_affine's early returns (c_api.cpp:1277-1285): no points
    writes nothing, a NULL array or head writes all-zero points."""
    import blst_hip
    L = blst_hip.load()
    dst = np.full((3, 12), 7, np.uint64)
    L.blst_p1s_to_affine(blst_hip._ptr(dst), None, 0)
    assert (dst == 7).all()
    L.blst_p1s_to_affine(blst_hip._ptr(dst), None, 3)
    assert not dst.any()
    dst[:] = 7
    heads = np.zeros(2, np.uint64)
    L.blst_p1s_to_affine(blst_hip._ptr(dst), blst_hip._ptr(heads), 3)
    assert not dst.any()


def test_fr_and_jacobian_generators(orc):
    """The Montgomery-form and Jacobian workload generators: blst_fr values
    equal k * 2^256 mod r (Python big integers) and convert back through the
    oracle; Jacobian representatives normalise back to their affine points."""
    import blst_hip
    sc = blst_hip.synth_scalars(500, 11)
    sc[0] = 0
    fr = blst_hip.synth_fr(sc)
    for s, f in zip(sc, fr):
        k = int.from_bytes(s.tobytes(), "little")
        assert orc.value(f) == (k << 256) % pyref.R_ORDER
    assert np.array_equal(orc.fr_to_scalars(fr), sc)
    aff = orc.gen_points(300, 12, threads=2)
    aff[::17] = 0
    jac = blst_hip.synth_jacobian(aff, 13)
    assert np.array_equal(orc.p1s_to_affine(jac), aff)
    assert all(j[12:].any() for j, a in zip(jac, aff) if a.any())
    assert not jac[::17].any()


def test_wbits_sizes_and_checks_without_gpu_work(orc):
    """Sizes of the fixed-base entry points against the oracle, and the
    argument checks of c_api.cpp:1318-1356 that return before any device call."""
    import blst_hip
    L = blst_hip.load()
    for wb in range(0, 17):
        for n in (0, 1, 5, 4096, 1 << 20, (1 << 64) - 1, (1 << 50)):
            assert L.blst_p1s_mult_wbits_precompute_sizeof(wb, n) == orc.wbits_precompute_sizeof(wb, n), (wb, n)
    for n in (0, 1, 100, 8191, 8192, 8193, 1 << 30):
        assert L.blst_p1s_mult_wbits_scratch_sizeof(n) == orc.wbits_scratch_sizeof(n)
    sc = np.zeros((2, 32), np.uint8)
    d = blst_hip.Descriptors(np.zeros((2, 12), np.uint64), sc)
    tab = np.zeros((4, 12), np.uint64)
    scr = np.zeros(64, np.uint64)
    ret = np.full(18, 7, np.uint64)
    P = blst_hip._ptr
    L.blst_p1s_mult_wbits(P(ret), P(tab), 2, 0, P(d.ss), 255, P(scr))
    assert not ret.any()
    for args in ((P(tab), 2, 2, P(d.ss), 0, P(scr)), (P(tab), 2, 2, P(d.ss), 257, P(scr)),
                 (P(tab), 1, 2, P(d.ss), 255, P(scr)), (P(tab), 15, 2, P(d.ss), 255, P(scr)),
                 (None, 2, 2, P(d.ss), 255, P(scr)), (P(tab), 2, 2, None, 255, P(scr)),
                 (P(tab), 2, 2, P(np.zeros(2, np.uint64)), 255, P(scr)), (P(tab), 2, 2, P(d.ss), 255, None)):
        ret[:] = 7
        L.blst_p1s_mult_wbits(P(ret), *args)
        assert not ret.any(), args
    # precompute: invalid wbits, NULL table or head write nothing
    pts = np.ones((2, 12), np.uint64)
    dp = blst_hip.Descriptors(pts, sc)
    tab[:] = 7
    for args in ((P(tab), 1, P(dp.pp), 2), (P(tab), 15, P(dp.pp), 2), (None, 2, P(dp.pp), 2),
                 (P(tab), 2, None, 2), (P(tab), 2, P(np.zeros(2, np.uint64)), 2), (P(tab), 2, P(dp.pp), 0)):
        L.blst_p1s_mult_wbits_precompute(*args)
        assert (tab == 7).all(), args


def test_g2_wbits_sizes_and_checks_without_gpu_work():
    """G2 fixed-base sizes: points_mult_wbits_precompute_sizeof<Fp2>
    (multi_scalar.h:244-256; sizeof(AffinePoint<Fp2>) = 192, zero for wbits
    outside 2..14 or on overflow) and points_mult_wbits_scratch_sizeof<Fp2>
    (multi_scalar.h:310-315: BatchAffineNode<Fp2>, an affine point and an Fp2,
    per point up to 4096, multi_scalar.h:218-222), with the node size from C
    layout rules; then the argument checks of c_api.cpp:1806-1888 that return
    before any device call."""
    import ctypes
    import blst_hip

    class Node(ctypes.Structure):
        _fields_ = [("point", ctypes.c_uint64 * 24), ("work", ctypes.c_uint64 * 12)]

    L = blst_hip.load()
    for wb in range(0, 17):
        for n in (0, 1, 5, 4096, 1 << 20, (1 << 64) - 1, 1 << 50, 1 << 56):
            want = 0
            if 2 <= wb <= 14 and 192 * n < (1 << 64) and (192 * n) << (wb - 1) < (1 << 64):
                want = (192 * n) << (wb - 1)
            assert L.blst_p2s_mult_wbits_precompute_sizeof(wb, n) == want, (wb, n)
    for n in (0, 1, 100, 4095, 4096, 4097, 1 << 30):
        assert L.blst_p2s_mult_wbits_scratch_sizeof(n) == ctypes.sizeof(Node) * min(n, 4096) == 288 * min(n, 4096)
    P = blst_hip._ptr
    sc = np.zeros((2, 32), np.uint8)
    ss = np.array([sc.ctypes.data, 0], np.uint64)
    tab = np.zeros((4, 24), np.uint64)
    scr = np.zeros(128, np.uint64)
    ret = np.full(36, 7, np.uint64)
    L.blst_p2s_mult_wbits(P(ret), P(tab), 2, 0, P(ss), 255, P(scr))
    assert not ret.any()
    for args in ((P(tab), 2, 2, P(ss), 0, P(scr)), (P(tab), 2, 2, P(ss), 257, P(scr)),
                 (P(tab), 1, 2, P(ss), 255, P(scr)), (P(tab), 15, 2, P(ss), 255, P(scr)),
                 (None, 2, 2, P(ss), 255, P(scr)), (P(tab), 2, 2, None, 255, P(scr)),
                 (P(tab), 2, 2, P(np.zeros(2, np.uint64)), 255, P(scr)), (P(tab), 2, 2, P(ss), 255, None)):
        ret[:] = 7
        L.blst_p2s_mult_wbits(P(ret), *args)
        assert not ret.any(), args
    pts = np.ones((2, 24), np.uint64)
    pp = np.array([pts.ctypes.data, 0], np.uint64)
    tab[:] = 7
    for args in ((P(tab), 1, P(pp), 2), (P(tab), 15, P(pp), 2), (None, 2, P(pp), 2),
                 (P(tab), 2, None, 2), (P(tab), 2, P(np.zeros(2, np.uint64)), 2), (P(tab), 2, P(pp), 0)):
        L.blst_p2s_mult_wbits_precompute(*args)
        assert (tab == 7).all(), args
    # batch normalisation: no points leaves dst alone, a NULL array or head
    # writes all-zero points (c_api.cpp:1806-1817)
    dst = np.full((3, 24), 7, np.uint64)
    L.blst_p2s_to_affine(P(dst), P(pp), 0)
    assert (dst == 7).all()
    for head in (None, P(np.zeros(2, np.uint64))):
        dst[:] = 7
        L.blst_p2s_to_affine(P(dst), head, 3)
        assert not dst.any()


def test_g2_scratch_sizes_and_checks(orc):
    """G2 scratch sizes (c_api.cpp:1890-1899: 384-byte Pointxyzz<Fp2>
    buckets) and the early returns that need no device."""
    import blst_hip
    L = blst_hip.load()
    for n in [0, 1, 2, 3, 17, 4096, 1 << 20]:
        assert L.blst_p2s_mult_pippenger_scratch_sizeof(n) == 2 * orc.scratch_sizeof(n)
    for w in range(0, 30):
        assert L.blst_p2s_mult_pippenger_window_scratch_sizeof(w) == 2 * orc.window_scratch_sizeof(w)
    ret = np.full(36, 7, np.uint64)
    P = blst_hip._ptr
    L.blst_p2s_mult_pippenger(P(ret), None, 0, None, 255, None)
    assert not ret.any()
    pts = np.zeros((2, 24), np.uint64)
    sc = np.zeros((2, 32), np.uint8)
    pp = np.array([pts.ctypes.data, 0], np.uint64)
    ss = np.array([sc.ctypes.data, 0], np.uint64)
    scr = np.zeros(4096, np.uint64)
    for args in ((P(pp), 2, P(ss), 0, P(scr)), (P(pp), 2, P(ss), 257, P(scr)), (P(pp), 2, P(ss), 255, None),
                 (P(np.zeros(2, np.uint64)), 2, P(ss), 255, P(scr))):
        ret[:] = 7
        L.blst_p2s_mult_pippenger(P(ret), *args)
        assert not ret.any(), args
    for w in (0, 25):
        ret[:] = 7
        L.blst_p2s_mult_pippenger_window(P(ret), P(pp), 2, P(ss), 255, P(scr), w)
        assert not ret.any()
        ret[:] = 7
        L.blst_p2s_tile_pippenger(P(ret), P(pp), 2, P(ss), 255, P(scr), 0, w)
        assert not ret.any()


def test_g2_point_generator(orc):
    """Synthetic G2 bases are an arithmetic progression of subgroup points:
    on the curve, P_{i+1} - P_i constant, [r] P = infinity."""
    import blst_hip
    import pyref2
    pts = blst_hip.synth_p2_points(6, 99)
    aff = [pyref2.limbs_to_affine(list(p)) for p in pts]
    assert all(pyref2.on_curve(a) for a in aff)
    d = pyref2.add(aff[1], pyref2.neg(aff[0]))
    for i in range(1, 5):
        assert pyref2.add(aff[i + 1], pyref2.neg(aff[i])) == d
    j = orc.p2_from_affine(pts[3])
    assert orc.p2_compress(orc.p2_mult(j, pyref2.R_ORDER.to_bytes(32, "little"), 256)) == bytes([0xC0]) + bytes(95)


def test_comb_table_bytes():
    import blst_hip
    L = blst_hip.load()
    assert L.blst_hip_comb_table_bytes(4096, 10, 255) == 26 * 4096 * 512 * 96
    assert L.blst_hip_comb_table_bytes(4096, 14, 255) == 19 * 4096 * 8192 * 96
    assert L.blst_hip_comb_table_bytes(7, 2, 1) == 1 * 7 * 2 * 96
    for args in ((0, 8, 255), (5, 1, 255), (5, 15, 255), (5, 8, 0), (5, 8, 257), ((1 << 62), 14, 255)):
        assert L.blst_hip_comb_table_bytes(*args) == 0, args


def test_tuning_rows_by_scalar_width():
    """The library's tuning lookup (host only): full-width scalars take the
    measured single-MSM rows, 33..64-bit and 65..128-bit scalars their own
    rows, other widths the cost model; every entry equals the committed
    table's line for that size."""
    import blst_hip
    rows = {}
    for line in open(os.path.join(ROOT, "temp-cat-blst_amd", "tuning.txt")):
        f = line.split("#")[0].split()
        if len(f) == 5:
            rows[(f[0], int(f[1]))] = tuple(int(x) for x in f[2:])
    for lg in (12, 16, 20):
        n = 1 << lg
        assert blst_hip.tuning_get(0, n, 255) == rows[("single", lg)]
        assert blst_hip.tuning_get(1, n, 255) == rows[("batch", lg)]
        assert blst_hip.tuning_get(0, n, 64) == rows[("single64", lg)]
        assert blst_hip.tuning_get(0, n, 40) == rows[("single64", lg)]
        assert blst_hip.tuning_get(0, n, 128) == rows[("single128", lg)]
        assert blst_hip.tuning_get(0, n, 100) == rows[("single128", lg)]
        # 1..32 and 129..199 bits: not tabulated, a window within the width's reach
        for nbits in (8, 32, 160):
            c, T, L = blst_hip.tuning_get(0, n, nbits)
            assert 1 <= c <= 24 and T >= 1 and L >= 1


def test_pipeline_settings_range():
    import blst_hip
    L = blst_hip.load()
    assert L.blst_hip_set_pipeline(17) == -1 and blst_hip.last_error()
    assert L.blst_hip_set_pipeline_share(51) == -1 and blst_hip.last_error()
    for v in (0, 1, 25, 50):
        assert L.blst_hip_set_pipeline_share(v) == 0
    assert L.blst_hip_set_pipeline(0) == 0 and L.blst_hip_set_pipeline_share(0) == 0
