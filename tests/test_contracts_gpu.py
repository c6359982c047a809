"""Contracts of the C ABI beyond the numbers: concurrent callers, failure of
device allocations, and the serving loop's admission limits.

The reference is reentrant and keeps no global state (README.md:39-40,
docs/API_CONTRACTS.md:175-179); its entry points fail closed, writing the
point at infinity (blst/c_api.cpp:1390-1408).  This backend keeps a device
context per GPU, so several host threads calling at once must each get their
own correct result, and an allocation failure must come back as infinity plus
an error message, never as an exception or a crash, with the next call
unaffected.  Allocation failures are forced from the host through the test
kit's allocation cap, so every kernel still runs to completion."""
import os
import subprocess
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
INF = bytes([0xC0]) + bytes(47)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def inputs(hip):
    pts = hip.synth_points((1 << 19) + 77, 0xC0C0)
    sc = hip.synth_scalars((1 << 19) + 77, 0xC0C1)
    return pts, sc


def test_concurrent_callers_each_get_their_result(hip, orc, inputs):
    """Four host threads at once: two plain MSMs (one large enough to run in
    point ranges), the blst_fr entry and G2; each result equals the oracle."""
    pts, sc = inputs
    g2 = hip.synth_p2_points(1000, 0xC0C2)
    g2s = hip.synth_scalars(1000, 0xC0C3)
    fr = hip.synth_fr(sc[:5000])
    want = {
        "small": orc.compress(orc.msm_pippenger(pts[:3000], sc[:3000], 3000, 255)),
        "ranged": orc.compress(orc.msm_pippenger(pts, sc, pts.shape[0], 255, threads=16)),
        "fr": orc.compress(orc.msm_pippenger(pts[:5000], sc[:5000], 5000, 255)),
        "g2": orc.p2_compress(orc.msm_p2_pippenger(g2, g2s, 1000, 255)),
    }
    calls = {
        "small": lambda: hip.compress(hip.p1s_mult_pippenger(pts[:3000], sc[:3000], 255)),
        "ranged": lambda: hip.compress(hip.p1s_mult_pippenger(pts, sc, 255)),
        "fr": lambda: (lambda ok_ret: hip.compress(ok_ret[1]) if ok_ret[0] else None)(
            hip.p1s_mult_fr(pts[:5000], fr, entry="msm")),
        "g2": lambda: orc.p2_compress(hip.p2s_mult_pippenger(g2, g2s, 255)),
    }
    got = {k: [] for k in calls}
    errors = []

    def run(name):
        try:
            for _ in range(3):
                got[name].append(calls[name]())
        except Exception as e:   # reported below, in the main thread
            errors.append((name, repr(e)))

    threads = [threading.Thread(target=run, args=(k,)) for k in calls]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in calls:
        assert got[k] == [want[k]] * 3, k


def test_allocation_failure_in_a_batch_plan(hip, orc, inputs):
    pts, sc = inputs
    n = 4096
    d_p, d_s = hip.DeviceBuffer(pts[:n]), hip.DeviceBuffer(sc[:n])
    jobs = [(d_p.ptr, d_s.ptr, n, 255)] * 3
    try:
        hip.set_alloc_cap(1 << 20)
        try:
            with pytest.raises(RuntimeError, match="allocation cap"):
                hip.BatchPlan(jobs)
        finally:
            hip.set_alloc_cap(0)
        plan = hip.BatchPlan(jobs)
        try:
            got = plan.run()
        finally:
            plan.close()
    finally:
        d_p.close()
        d_s.close()
    want = orc.compress(orc.msm_pippenger(pts[:n], sc[:n], n, 255))
    assert [hip.compress(g) for g in got] == [want] * 3


def test_allocation_failure_in_the_serving_loop(hip, orc, inputs):
    """The copier cannot get device memory for a fresh-base request: every
    admitted request fails with an error text; a later server serves."""
    pts, sc = inputs
    n = 4096
    hip.set_alloc_cap(1 << 20)
    try:
        srv = hip.Server()
        try:
            t = srv.submit(sc[:n], 255, points=pts[:n])
            with pytest.raises(RuntimeError, match="request failed"):
                srv.wait(t)
        finally:
            srv.close()
    finally:
        hip.set_alloc_cap(0)
    srv = hip.Server()
    try:
        got, _ = srv.wait(srv.submit(sc[:n], 255, points=pts[:n]))
    finally:
        srv.close()
    assert hip.compress(got) == orc.compress(orc.msm_pippenger(pts[:n], sc[:n], n, 255))


_CHILD = r"""
import sys
sys.path[:0] = [sys.argv[1] + "/temp-cat-blst_amd", sys.argv[1] + "/oracle"]
import numpy as np
import blst_hip as hip
pts = np.load(sys.argv[2]); sc = np.load(sys.argv[3])
L = hip.load()
n = pts.shape[0]
d = hip.Descriptors(pts, sc, "contiguous")
scr = np.zeros(hip.scratch_sizeof(n) // 8, np.uint64)


def call():
    ret = np.full(18, 0x5555, np.uint64)        # not infinity before the call
    L.blst_p1s_mult_pippenger(hip._ptr(ret), hip._ptr(d.pp), n, hip._ptr(d.ss), 255, hip._ptr(scr))
    return hip.compress(ret).hex(), hip.last_error()


hip.set_alloc_cap(1 << 20)
first, err = call()
hip.set_alloc_cap(0)
second, err2 = call()
assert err2 == "", err2
print(first, second, err.replace(" ", "_"))
"""


def test_allocation_failure_in_the_drop_in_entry(hip, orc, inputs, tmp_path):
    """In a fresh process (the first call is the one that allocates the
    context's buffers): blst_p1s_mult_pippenger under the cap writes infinity
    and sets blst_hip_last_error(); the next call, without the cap, is
    correct."""
    pts, sc = inputs
    n = 5000
    fp, fs = str(tmp_path / "p.npy"), str(tmp_path / "s.npy")
    np.save(fp, pts[:n])
    np.save(fs, sc[:n])
    print("child process: drop-in entry under an allocation cap", flush=True)
    out = subprocess.run([sys.executable, "-c", _CHILD, ROOT, fp, fs], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    first, second, err = out.stdout.split()[-3:]
    assert bytes.fromhex(first) == INF
    assert "allocation_cap" in err
    assert bytes.fromhex(second) == orc.compress(orc.msm_pippenger(pts[:n], sc[:n], n, 255))


def test_staging_failure_in_a_split_host_msm(hip, orc, inputs):
    """A host-buffer MSM of 2^19 + 77 points runs as point ranges staged by a
    thread of the library's own while the caller's thread queues each range's
    kernels.  With that thread made to fail before its second range, the call
    writes infinity and sets blst_hip_last_error(); the next call is correct."""
    pts, sc = inputs
    n = pts.shape[0]
    hip.set_stage_failure(1)
    try:
        print("drop-in MSM of %d points with the staging of range 1 refused" % n, flush=True)
        with pytest.raises(RuntimeError, match="staging of point range 1 refused"):
            hip.p1s_mult_pippenger(pts, sc, 255)
    finally:
        hip.set_stage_failure(-1)
    got = hip.compress(hip.p1s_mult_pippenger(pts, sc, 255))
    assert got == orc.compress(orc.msm_pippenger(pts, sc, n, 255, threads=16))


def test_server_rejects_what_it_cannot_run(hip, inputs):
    pts, sc = inputs
    srv = hip.Server(hbm_budget=1 << 20)
    try:
        with pytest.raises(RuntimeError, match="larger than the HBM budget"):
            srv.submit(sc[:4096], 255, points=pts[:4096])
        sid = srv.add_bases(pts[:100])
        with pytest.raises(RuntimeError, match="more points than the bases hold"):
            srv.submit(sc[:200], 255, base_id=sid, n=200)
    finally:
        srv.close()
    srv = hip.Server()
    try:
        # 2^28 points of 255-bit scalars are more (window, point) entries than
        # one launch indexes: rejected before anything is read or copied
        with pytest.raises(RuntimeError, match="too large for one launch"):
            srv.submit(sc[:16], 255, points=pts[:16], n=1 << 28)
        got, _ = srv.wait(srv.submit(sc[:64], 255, points=pts[:64]))   # still serving
        assert hip.compress(got) != bytes(48)
    finally:
        srv.close()


def test_server_ticket_waited_from_two_threads(hip, orc, inputs):
    """Two threads waiting for one ticket both get the result (the ticket is
    forgotten by key once, not erased twice); a later wait is refused."""
    pts, sc = inputs
    want = orc.compress(orc.msm_pippenger(pts, sc, pts.shape[0], 255, threads=16))
    srv = hip.Server()
    try:
        t = srv.submit(sc, 255, points=pts)   # 2^19 + 77 fresh points: both threads wait well before it ends
        got, errors = [], []

        def waiter():
            try:
                got.append(hip.compress(srv.wait(t)[0]))
            except RuntimeError as e:         # the other thread forgot the ticket first
                errors.append(str(e))

        threads = [threading.Thread(target=waiter) for _ in range(2)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        assert got and all(g == want for g in got), (len(got), errors)
        assert all("unknown ticket" in e for e in errors), errors
        with pytest.raises(RuntimeError, match="unknown ticket"):
            srv.wait(t)
        got2, _ = srv.wait(srv.submit(sc[:64], 255, points=pts[:64]))   # still serving
        assert hip.compress(got2) == orc.compress(orc.msm_pippenger(pts[:64], sc[:64], 64, 255))
    finally:
        srv.close()


def test_batch_beyond_one_dispatch_is_refused(hip, orc, inputs):
    """A batch whose launches would pass 2^32 work-items (32,768 jobs of 64
    points in 256 one-bit windows: 2^23 sort tiles of 512 threads) is refused
    before anything runs, with the reason; the next batch runs normally."""
    pts, sc = inputs
    dp, ds = hip.DeviceBuffer(pts[:64]), hip.DeviceBuffer(sc[:64])
    try:
        with pytest.raises(RuntimeError, match="too large for one launch"):
            hip.msm_batch_device([(dp.ptr, ds.ptr, 64, 255, 1)] * 32768)
        got = hip.msm_batch_device([(dp.ptr, ds.ptr, 64, 255, 1)] * 3)
        want = orc.compress(orc.msm_pippenger(pts[:64], sc[:64], 64, 255))
        assert all(hip.compress(g) == want for g in got)
    finally:
        dp.close()
        ds.close()
