import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU test (full-size grids), run explicitly with -m slow")


def pytest_collection_modifyitems(config, items):
    # slow tests only run when selected explicitly with -m slow
    expr = config.getoption("-m") or ""
    if "slow" in expr:
        return
    skip = pytest.mark.skip(reason="full-size CPU test; select with -m slow")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def wx():
    import woxel_amd

    return woxel_amd


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    return torch
