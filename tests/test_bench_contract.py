"""bench.py's contract on a machine without a GPU: more than one GPU is
refused with a message, and the workload generators are deterministic and
within their stated ranges."""
import os
import subprocess
import sys

import numpy as np

from mplb_amd import cases
from mplb_amd import warmstart as ws

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_batch_spec_ranges_and_determinism():
    sizes, res = cases.batch_spec(256)
    assert set(sizes) <= set(cases.BATCH_SIZES)
    assert len(set(sizes)) == len(cases.BATCH_SIZES)  # every size drawn
    assert 100.0 <= min(res) and max(res) <= 1000.0
    assert (sizes, res) == cases.batch_spec(256)
    nodes = sum(n * n for n in sizes)
    assert 6.0e6 < nodes < 8.5e6  # E[N^2] = 26,419 per case


def test_warmstart_cases_are_valid_obstacle_layouts():
    res, rects = ws.case_params(64, seed=3, n=128)
    assert all(100.0 <= r <= 1000.0 for r in res)
    for rs in rects:
        assert len(rs) <= 3
        for x0, y0, w, h in rs:
            assert 8 <= w <= 32 and 8 <= h <= 32
            assert x0 >= 2 and y0 >= 2 and x0 + w <= 126 and y0 + h <= 126
    x = ws.inputs(128, res[:4], rects[:4])
    assert x.shape == (4, 128, 128, ws.IN_CH)
    assert np.array_equal(x[..., 0], np.stack([cases.rect_mask(128, r) for r in rects[:4]]))
    assert np.all(x[:, 127, :, 3] == 1) and np.all(x[:, :127, :, 3] == 0)
