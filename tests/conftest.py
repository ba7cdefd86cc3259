import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, 'sporco-cuda_amd'), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an MI355X (runs through the HIP C ABI)')


@pytest.fixture(scope='session')
def gpu():
    """Skip unless a HIP device is visible; the HIP extension must load."""
    import sporco_cuda.util as cu
    if cu.device_count() < 1:
        pytest.skip('no HIP device')
    return 0
