import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "raster-join_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

ORACLE_LIB = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
RJ_LIB = os.path.join(PKG, "lib", "librasterjoin.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    # Build what is missing (the GPU box receives prebuilt libraries).
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    if not os.path.exists(RJ_LIB):
        subprocess.run(["make", "-C", PKG, "-j8"], check=True, stdout=subprocess.DEVNULL)


_gpu = None


def gpu_present():
    global _gpu
    if _gpu is None:
        import rasterjoin as rj
        try:
            h = rj.Handler()
            h.close()
            _gpu = True
        except rj.RJError:
            _gpu = False
    return _gpu


@pytest.fixture
def handler():
    import rasterjoin as rj
    if not gpu_present():
        pytest.fail("GPU test run without a visible GPU")
    h = rj.Handler()
    yield h
    h.close()
