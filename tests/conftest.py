import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X) and the built library')
    config.addinivalue_line('markers', 'slow: long-running')


@pytest.fixture(scope='session')
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from opencmp_amd.core import Context
    return Context.default(0)
