import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")
    # a state that goes non-finite fails a test unless the test expects it (pytest.warns)
    config.addinivalue_line("filterwarnings", "error::fv3d_amd.NonFiniteWarning")


@pytest.fixture(scope="session")
def gpu_available():
    import fv3d_amd
    fv3d_amd.lib()  # raises if libfv3d_hip.so is missing: no silent fallback
    return True
