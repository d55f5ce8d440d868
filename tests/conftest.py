import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X) and the built HIP library')
    config.addinivalue_line('markers', 'slow: long-running statistical test')


def pytest_collection_modifyitems(config, items):
    # run GPU tests in collection order, but keep the no-allocation entry point test last
    pass


@pytest.fixture(scope='session')
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail('GPU test collected on a machine without a ROCm GPU')
    from covasim_amd import _lib
    _lib.lib()
    return torch.device('cuda')
