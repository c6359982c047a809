"""blst_hip_msm_device_noalloc (no counterpart in the reference): the MSM of
blst_hip_msm_device computed in the library's fixed per-device scratch array,
for callers that cannot allocate.  Through the C ABI, its result must equal
blst_hip_msm_device's on the same device inputs, byte for byte, on calls in a
row whose sizes go down (the array keeps the previous call's contents), at
the largest size it accepts, and for short scalars; beyond 65536 points it
refuses with an error and writes infinity."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MAX_POINTS = 1 << 16
INF = bytes([0xC0]) + bytes(47)


def _noalloc(hip, d_p, d_s, n, nbits):
    ret = np.full(18, 0x77, np.uint64)
    rc = hip.load().blst_hip_msm_device_noalloc(hip._ptr(ret), ctypes.c_void_p(d_p), ctypes.c_void_p(d_s), n, nbits)
    return rc, ret


def _plain(hip, d_p, d_s, n, nbits):
    ret = np.zeros(18, np.uint64)
    rc = hip.load().blst_hip_msm_device(hip._ptr(ret), ctypes.c_void_p(d_p), ctypes.c_void_p(d_s), n, nbits, 0)
    assert rc == 0, hip.last_error()
    return ret


@pytest.fixture(scope="module")
def device_inputs(hip):
    n = MAX_POINTS + 64
    pts = hip.synth_points(n, 0xA110C)
    sc = hip.synth_scalars(n, 0xA110D)
    hip.synth_edge_cases(pts, sc, 0xA110E)
    d_p, d_s = hip.DeviceBuffer(pts), hip.DeviceBuffer(sc)
    yield d_p, d_s, pts, sc
    d_p.close()
    d_s.close()


def test_noalloc_equals_device_msm_on_calls_in_a_row(hip, orc, device_inputs):
    d_p, d_s, pts, sc = device_inputs
    # the largest accepted size first, then smaller ones: every later call
    # runs over the array the earlier ones left behind
    for n in (MAX_POINTS, 40000, 4096, 777, 3, 1):
        rc, got = _noalloc(hip, d_p.ptr, d_s.ptr, n, 255)
        assert rc == 0, (n, hip.last_error())
        assert hip.compress(got) == hip.compress(_plain(hip, d_p.ptr, d_s.ptr, n, 255)), n
    want = orc.compress(orc.msm_pippenger(pts[:777], sc[:777], 777, 255))
    assert hip.compress(_noalloc(hip, d_p.ptr, d_s.ptr, 777, 255)[1]) == want


def test_noalloc_every_tuned_size_class_fits(hip, device_inputs):
    """The largest size of every size class and scalar widths of 256, 128, 64
    and 1 bits (the window and tile, and so the scratch a call needs, follow
    the class and the width's row of the tuning table)."""
    d_p, d_s, _, _ = device_inputs
    for k in range(16, 4, -1):
        n = min((2 << k) - 1, MAX_POINTS)
        for nbits in (256, 128, 64, 1):
            rc, got = _noalloc(hip, d_p.ptr, d_s.ptr, n, nbits)
            assert rc == 0, (n, nbits, hip.last_error())
            assert hip.compress(got) == hip.compress(_plain(hip, d_p.ptr, d_s.ptr, n, nbits)), (n, nbits)


def test_noalloc_limits(hip, device_inputs):
    d_p, d_s, _, _ = device_inputs
    rc, got = _noalloc(hip, d_p.ptr, d_s.ptr, MAX_POINTS + 1, 255)
    assert rc == -1 and hip.compress(got) == INF
    assert "65536-point limit" in hip.last_error()
    rc, got = _noalloc(hip, d_p.ptr, d_s.ptr, 0, 255)
    assert rc == 0 and hip.compress(got) == INF
    rc, got = _noalloc(hip, d_p.ptr, d_s.ptr, 100, 257)
    assert rc == -1 and hip.compress(got) == INF
    # still correct after the refusals
    rc, got = _noalloc(hip, d_p.ptr, d_s.ptr, 1000, 255)
    assert rc == 0 and hip.compress(got) == hip.compress(_plain(hip, d_p.ptr, d_s.ptr, 1000, 255))
