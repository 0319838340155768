import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Every search launch stops itself after this many seconds instead of running
# on (device watchdog); GPU tests are sized to finish far below it.
os.environ.setdefault("TDP_KERNEL_TIMEOUT_S", "60")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; a GPU test fails (never skips) when it is missing."""
    import tdp1375_amd
    lib = tdp1375_amd.load()
    if lib.tdp_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return lib


@pytest.fixture(scope="module")
def slots(gpu_lib):
    """Two slots on device 0: 2 GiB of queue memory, 2^22 dedup entries."""
    import tdp1375_amd
    tdp1375_amd.init_slots(2, 1, 2 << 30, 1 << 22)
    yield [0, 1]
    tdp1375_amd.destroy_slots()
