import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(params=["split-fold0", "split-fold1", "split-fold2", "split-fold3", "fused"])
def step_form(request, monkeypatch):
    """Run a test with every form of the step: bulk + boundary kernels with
    the bulk kernel running none, the x-face, all simple boundary regions of
    the planes z = 1 .. n-2, or every simple region of every plane (3D; 2D
    cases take level 2 for 3) -- the forms cases of 2^18 nodes and more
    choose between -- and the single whole-lattice launch (what smaller cases
    use).  The library reads these variables when a case is prepared."""
    fused = request.param == "fused"
    monkeypatch.setenv("MPLBHIP_FUSED", "1" if fused else "0")
    if not fused:
        monkeypatch.setenv("MPLBHIP_XFACE_FOLD", request.param[-1])
    return request.param
