import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "era-bellman-cuda_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through its C ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    if not os.path.exists(O.LIB_PATH):
        import subprocess

        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    O.lib()
    return O


@pytest.fixture(scope="session")
def bellman():
    """The HIP library with ff/ntt/pn/msm set up on device 0 (GPU tests only)."""
    import ctypes

    import bellman_amd as B

    B.call("ff_set_up", ctypes.c_uint(25), ctypes.c_uint(14))
    B.call("ntt_set_up")
    B.call("pn_set_up")
    B.call("msm_set_up")
    dev = B.Device(0)
    yield B, dev
    dev.close()
