import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP) device')


def pytest_collection_modifyitems(config, items):
    # GPU tests are only meaningful where a HIP device exists; a CPU-only run
    # selects them out with -m "not gpu".
    pass


@pytest.fixture(scope='session')
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail('GPU test run without a HIP device')
    return torch.device('cuda', 0)
