import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def hip():
    """The HIP library; on a GPU test run a missing library or device is an error."""
    from openmm_rigidbody_plugin_amd import rbhip
    rbhip.lib()
    n = rbhip.device_count()
    if n < 1:
        raise RuntimeError("no GPU visible to the HIP library")
    return rbhip
