"""The drop-in boundary from a plain C11 program (tests/c_caller.c): built
with gcc -std=c11 -pedantic -Wall -Wextra -Werror against include/blst_hip.h
and linked with libblst_hip.so, as a C caller of the reference library
would be.  Without a GPU it checks the scratch-size companions and the
argument checks that return before any device work; on the GPU it runs an
MSM through blst_p1s_mult_pippenger and compares the compressed result with
the oracle's (linked into the program as test infrastructure)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "temp-cat-blst_amd")
FLAGS = ["-std=c11", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1"]


def _build(tmp, oracle):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    if not os.path.exists(os.path.join(PKG, "libblst_hip.so")):
        raise RuntimeError("libblst_hip.so is not built (run __graft_entry__.build())")
    exe = os.path.join(tmp, "c_caller")
    cmd = ["gcc"] + FLAGS + ["-o", exe, os.path.join(HERE, "c_caller.c")]
    if oracle:
        cmd += ["-DWITH_ORACLE", "-Wno-pedantic", os.path.join(ROOT, "oracle", "blst_oracle.c"),
                os.path.join(ROOT, "oracle", "blst_oracle_g2.c"), "-pthread"]
    cmd += ["-L" + PKG, "-l:libblst_hip.so", "-Wl,-rpath," + PKG]
    subprocess.run(cmd, check=True)
    return exe


def test_c_caller_checks_without_gpu_work(tmp_path):
    exe = _build(str(tmp_path), oracle=False)
    p = subprocess.run([exe, "checks"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and p.stdout.strip() == "ok", p.stdout + p.stderr


@pytest.mark.gpu
def test_c_caller_msm_matches_oracle(hip, tmp_path):
    exe = _build(str(tmp_path), oracle=True)
    for n in ("1", "777", "70000"):
        p = subprocess.run([exe, "msm", n], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0 and p.stdout.strip() == "ok", (n, p.stdout + p.stderr)
