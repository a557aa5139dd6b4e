import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: longer CPU run (tens of seconds)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import samurai_amd as sa
    ctx = sa.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def host_ctx():
    import samurai_amd as sa
    ctx = sa.Context(-1)
    yield ctx
    ctx.close()
