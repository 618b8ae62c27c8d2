import os
import sys

import pytest

# One HIP runtime per test process.  torch ships its own copies of the HIP
# and HSA runtime libraries under the same sonames as the system ROCm's
# (libamdhip64.so.7, libhsa-runtime64.so.1).  Loaded before libsmc.so, they
# are the ones libsmc.so binds to, so a test that hands torch device memory
# to the library shares a single runtime with it; loaded after, torch brings
# a second runtime that no longer sees the GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle
