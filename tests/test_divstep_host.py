"""The divstep inversion of csrc/fp_inv.hip.h, built for the host with g++:
inverses modulo p against Python's pow(x, -1, p) on the values where a
binary gcd goes wrong first (0, 1, p - 1, powers of two, p - 2^k, 2^k - 1)
and on random values.  The device build of the same header is checked by the
batch-normalisation tests on the GPU."""
import os
import random
import shutil
import subprocess

import pytest

import pyref

P = pyref.P
HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "temp-cat-blst_amd", "csrc", "fp_inv.hip.h")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#define BLS_HD inline
#include HEADER_PATH
int main() {
    char line[256];
    while (fgets(line, sizeof line, stdin)) {
        uint32_t x[12], out[12];
        for (int w = 0; w < 12; w++) {
            char buf[9];
            memcpy(buf, line + (11 - w) * 8, 8);
            buf[8] = 0;
            x[w] = (uint32_t)strtoul(buf, nullptr, 16);
        }
        bls::divstep::inverse(out, x);
        for (int w = 11; w >= 0; w--) printf("%08x", out[w]);
        printf("\n");
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("divstep")
    src = d / "drv.cpp"
    src.write_text(DRIVER.replace("HEADER_PATH", '"%s"' % HEADER))
    exe = d / "drv"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-unknown-pragmas", "-o", str(exe), str(src)], check=True)
    return str(exe)


def _run(exe, xs):
    inp = "".join("%096x\n" % x for x in xs)
    out = subprocess.run([exe], input=inp.encode(), capture_output=True, check=True).stdout.decode().split()
    assert len(out) == len(xs)
    return [int(o, 16) for o in out]


def test_inverse_edge_values(driver):
    xs = [0, 1, 2, 3, P - 1, P - 2, (P - 1) // 2, (P + 1) // 2]
    xs += [1 << k for k in range(381)]
    xs += [P - (1 << k) for k in range(381)]
    xs += [(1 << k) - 1 for k in range(1, 381)]
    for x, v in zip(xs, _run(driver, xs)):
        assert v < 2 * P, hex(x)
        assert v % P == (pow(x, -1, P) if x else 0), hex(x)


def test_inverse_random(driver):
    rnd = random.Random(0x1D5)
    xs = [rnd.randrange(1, P) for _ in range(20000)]
    xs += [rnd.getrandbits(rnd.randrange(1, 381)) or 1 for _ in range(5000)]   # short values
    for x, v in zip(xs, _run(driver, xs)):
        assert v < 2 * P and v * x % P == 1, hex(x)
