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


@pytest.fixture(params=[(0, 64), (3, 64), (5, 64), (0, 32), (3, 32)],
                ids=["nzc_auto", "nzc3", "nzc5", "w32", "w32_nzc3"])
def nzc(request, monkeypatch):
    """Launch shape of the stage kernel for contexts made in the test.
    z-chunk length: the library's own choice (one plane per workgroup on
    small boxes), or short chunks of 3 and 5 planes, so each workgroup
    marches several planes -- carrying z slopes, faces, the temperature ring
    and the viscous planes from plane to plane -- and the last chunk of most
    boxes is partial.  Tile width: 64 columns (one window row per wave) or
    32 (two rows per wave, the halo columns shared by the two halo waves)."""
    n, wx = request.param
    if n:
        monkeypatch.setenv("FV3D_NZC", str(n))
    monkeypatch.setenv("FV3D_TILE_WX", str(wx))
    return request.param


@pytest.fixture(params=[64, 32], ids=["w64", "w32"])
def tile_wx(request, monkeypatch):
    """Stage tile width (columns) for contexts and ensembles made in the test."""
    monkeypatch.setenv("FV3D_TILE_WX", str(request.param))
    return request.param


@pytest.fixture(scope="session")
def gpu_available():
    import fv3d_amd
    fv3d_amd.lib()  # raises if libfv3d_hip.so is missing: no silent fallback
    return True
