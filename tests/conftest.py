import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libbinder_hip.so")


@pytest.fixture(scope="session")
def orc():
    import pycore
    return pycore


@pytest.fixture(scope="session")
def core():
    from binder_amd import _core
    return _core


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
