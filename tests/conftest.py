"""Test set-up: the product package and the oracle (the checker) on sys.path,
and the ``gpu`` marker for tests that need an MI355X."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "spin-tracking_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


def to_oracle_opt(opt):
    """product Options (ctypes) -> oracle options record (same 440 bytes)."""
    import oracle as O
    return np.frombuffer(bytes(opt), dtype=O.OPTIONS_DTYPE).copy().reshape(())
