import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (GPU box only)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(params=["split", "fused"])
def step_form(request, monkeypatch):
    """Run a test with both forms of the step: bulk + boundary kernels
    (what cases of 2^18 nodes and more use) and the single whole-lattice
    launch (what smaller cases use).  The library reads MPLBHIP_FUSED when
    a case is prepared."""
    monkeypatch.setenv("MPLBHIP_FUSED", "1" if request.param == "fused" else "0")
    return request.param
