import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnumc_amd.so on the GPU)")


def _ensure_built():
    import numc_amd
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    if not os.path.exists(numc_amd.LIB_PATH):
        from numc_amd import build as b
        b.build()


_ensure_built()


@pytest.fixture
def ctx():
    import numc_amd
    c = numc_amd.Ctx()
    yield c
    c.close()
