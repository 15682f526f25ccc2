import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP path through the C-ABI")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle, built from oracle/ if needed (checker only)."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    from oracle import oracle as O
    return O


@pytest.fixture
def oracle(oracle_lib):
    o = oracle_lib.Oracle()
    yield o
    o.close()


@pytest.fixture(scope="session")
def gpu_ctx():
    from tetra3d_amd.runtime import Context
    c = Context(0)
    # every render starts from NaN-filled outputs, so a vertex the path fails
    # to write cannot pass by keeping an earlier test's value
    c.poison_outputs(0x7FC0DEAD)
    yield c
    c.close()
