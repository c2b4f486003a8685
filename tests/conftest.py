"""pytest configuration: the `gpu` marker, import paths and session fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs on the MI355X)")
    config.addinivalue_line("markers", "slow: long-running (large sizes)")


def _ensure_built():
    from vulkpy_amd import _lib
    from oracle import oracle as O
    if not os.path.exists(_lib.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    if not os.path.exists(O.LIB_PATH):
        O.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import vulkpy_amd as vk
    return vk.GPU()


@pytest.fixture
def rng():
    import numpy as np
    return np.random.default_rng(0)
