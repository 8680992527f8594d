import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


# Iteration bound of every planner run in the GPU tests: a kernel that broke
# the tree's growth makes its test fail instead of running without end.
GPU_ITER_CAP = 200_000


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle.lib()


@pytest.fixture(scope="session")
def ctx_factory():
    """Creates GPU contexts; every context is closed at the end of the session."""
    import mpt_amd
    made = []

    def make():
        c = mpt_amd.Context(0)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()
