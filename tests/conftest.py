import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(ROOT, "oracle")
for d in (TESTS, ORACLE):
    if d not in sys.path:
        sys.path.insert(0, d)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running oracle comparison")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session."""
    import ratbite_amd
    c = ratbite_amd.Context(0)
    yield c
    c.close()
