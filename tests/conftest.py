import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "temp-cat-blst_amd"), os.path.join(ROOT, "oracle"),
          os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The product library on a live GPU; GPU tests fail (not skip) without one."""
    import blst_hip
    blst_hip.load()
    if not blst_hip.available():
        raise RuntimeError("gpu test without a visible gfx950 device")
    return blst_hip
