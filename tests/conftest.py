import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def cx_runtime():
    """The HIP product; on a machine with a GPU a missing library is an error, never a skip."""
    from crisoptix_amd import runtime as rt
    rt.lib()
    if rt.device_count() < 1:
        pytest.fail("no HIP device visible to a test marked gpu")
    return rt


@pytest.fixture(scope="session")
def cx_ctx(cx_runtime):
    return cx_runtime.Context(0)
