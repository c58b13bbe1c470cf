import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "variational-monte-carlo_amd", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def vmc():
    import vmc_amd

    vmc_amd.device_check()
    return vmc_amd


@pytest.fixture(scope="session")
def vmc_nodev():
    """The binding without a device check (CPU-side C ABI tests)."""
    import vmc_amd

    vmc_amd.lib()
    return vmc_amd
