"""Shared test configuration: the `gpu` marker and numeric helpers."""
import os
import sys

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels through the C-ABI)')


def pytest_collection_modifyitems(config, items):
    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason='no GPU in this container')
    for item in items:
        if 'gpu' in item.keywords:
            item.add_marker(skip)


def rel_err(a, b):
    """Norm-wise relative error ||a - b|| / ||b|| (float64)."""
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    nb = b.norm().item()
    d = (a - b).norm().item()
    return d / nb if nb > 0 else d


@pytest.fixture
def dev():
    return torch.device('cuda:0')
