"""Host code under sanitizers (CPU only).

tests/host_sanitize.cpp checks the product's host code (the G1 and G2 window
folds, compression, batch normalisation and inversion of csrc/host_ec.h and
csrc/host_ec2.h, the 28-bit limb field of csrc/fq.hip.h with its bucket sum,
the divstep inversion of csrc/fp_inv.hip.h) against the oracle.  It is built
here with g++ and the oracle's C sources:

* with AddressSanitizer and UndefinedBehaviorSanitizer, every check fatal
  (-fno-sanitize-recover=all), for the single-threaded code;
* with ThreadSanitizer, for the oracle's threaded MSMs, batches and point
  generation against their one-thread forms.

The device builds of the same headers are checked on the GPU by the parity
tests; the product's host threads (staging pool, serving loop) call the HIP
runtime and are covered by the GPU contract tests instead.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE_C = [os.path.join(ROOT, "oracle", f) for f in ("blst_oracle.c", "blst_oracle_g2.c")]
DRIVER = os.path.join(HERE, "host_sanitize.cpp")

MODES = {
    "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
    "tsan": ["-fsanitize=thread"],
}


def _build(tmp, mode):
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("no host C/C++ compiler")
    flags = MODES[mode] + ["-g", "-O1", "-fno-omit-frame-pointer"]
    objs = []
    for src in ORACLE_C:
        obj = os.path.join(tmp, os.path.basename(src) + ".o")
        subprocess.run(["gcc", "-std=gnu11", "-c", "-o", obj, src] + flags, check=True)
        objs.append(obj)
    exe = os.path.join(tmp, "host_sanitize_" + mode)
    subprocess.run(["g++", "-std=c++17", "-Wno-unknown-pragmas", "-o", exe, DRIVER] + objs + flags + ["-pthread"],
                   check=True)
    return exe


def _run(exe, mode):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:second_deadlock_stack=1"
    p = subprocess.run([exe, mode], capture_output=True, text=True, env=env, timeout=600)
    assert p.returncode == 0, "%s\n%s" % (p.stdout[-4000:], p.stderr[-8000:])
    assert p.stdout.startswith("ok "), p.stdout
    assert "runtime error" not in p.stderr and "WARNING: ThreadSanitizer" not in p.stderr, p.stderr[-8000:]
    return int(p.stdout.split()[1])


def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = _build(str(tmp_path), "asan")
    assert _run(exe, "asan") > 1000


def test_oracle_threads_under_tsan(tmp_path):
    exe = _build(str(tmp_path), "tsan")
    assert _run(exe, "tsan") == 8
