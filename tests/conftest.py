"""pytest configuration: markers, import paths and the shared device fixture.

GPU tests are marked ``gpu``.  When no HIP device is visible they are skipped
with a message (the reference's tests likewise exit early when no SYCL device
exists, tests/include/test_utils.hpp:49-61); a missing libsygraph_hip.so is an
error, never a skip.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: large-graph test (minutes on the CPU oracle)")


@pytest.fixture(scope="session")
def sg():
    import sygraph_amd

    sygraph_amd.lib()  # raises ImportError when the library was not built
    return sygraph_amd


@pytest.fixture(scope="session")
def device(sg):
    if sg.device_count() == 0:
        pytest.skip("no HIP device visible")
    d = sg.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle
