import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def engine():
    """The HIP engine; fails loudly (no fallback) when it is not built."""
    import rtdbstream_amd as R
    R.lib()
    if R.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return R
