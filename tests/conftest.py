import os
import sys

import pytest

# torch ships its own HIP runtime under the same soname as /opt/rocm's; the
# first one loaded serves the whole process, and torch only initialises its
# GPU support on its own.  Tests that hand torch device buffers to the
# library therefore load torch before libasv.so.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def have_gpu():
    import anisotropic_supervoxel_amd as asv
    n = asv._lib.load().asv_device_count()
    if n < 1:
        pytest.skip("no HIP device")
    return True
