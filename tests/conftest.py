import os
import sys

import pytest

# torch ships its own copy of the HIP runtime (same soname as /opt/rocm's).
# Loaded first, it is the one libcubnm_amd.so binds to; loaded after
# libcubnm_amd.so it would be a second runtime in the process and torch would
# find no GPU.  Tests that hand torch tensors to the library therefore need
# torch imported before anything loads the library.
import torch  # noqa: F401,E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.lib()
    return O
