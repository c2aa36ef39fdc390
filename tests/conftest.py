import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible AMD GPU (gfx950) and the built libpado_gpu.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def gpu_ctx():
    import parallel_distance_oracle_amd as pado

    if pado.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    ctx = pado.Context(0)
    yield ctx
    ctx.close()
