import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="module")
def hip():
    """The HIP backend on device 0, vectors in LDS wherever they fit (the reference planner's rule)."""
    from montmul_amd import Hip

    h = Hip(prefer_lds=True)
    yield h
    h.close()


@pytest.fixture(scope="module")
def hip_global():
    """The HIP backend with global mode forced (GpuOptions::force_global)."""
    from montmul_amd import Hip

    h = Hip(force_global=True)
    yield h
    h.close()
