import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X)')


def pytest_runtest_setup(item):
    if 'gpu' in item.keywords:
        import torch
        if not torch.cuda.is_available():
            pytest.skip('no GPU')
