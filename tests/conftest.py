import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test of a minute or more")


def pytest_collection_modifyitems(config, items):
    # a gpu test on a machine without a GPU is skipped (the driver runs them on the MI355X box)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU here")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
