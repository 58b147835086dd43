import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def ssc():
    """The product package with its HIP library loaded; GPU tests fail loudly without it."""
    import torch
    import streamsculptor_amd as S
    from streamsculptor_amd import _lib
    assert torch.cuda.is_available(), "GPU test collected without a visible ROCm GPU"
    _lib.require_gpu()
    return S
