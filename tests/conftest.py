import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def grrt():
    import grrt_amd
    grrt_amd.lib()
    if not grrt_amd.gpu_available():
        pytest.fail("no HIP device visible to libgrrt.so: GPU tests need an MI355X")
    return grrt_amd
