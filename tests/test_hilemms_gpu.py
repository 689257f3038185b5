"""The HiLeMMS C++ drivers on the GPU against the CPU oracle: a JSON case
read by the C++ layer and run through DefineCase ... Iterate must reproduce
the oracle's macroscopic fields bit for bit (FP64 storage)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from mplb_amd import build as hipbuild
from mplb_amd import cases
from oracle.oracle import OracleCase

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(exe, *args):
    print("run", os.path.basename(exe), *args, flush=True)
    r = subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, MPLB_STORAGE="f64"))
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def split_dump(path, sizes, nvars):
    data = np.fromfile(path, dtype=np.float64)
    out, p = {}, 0
    for b in sorted(sizes):
        shp = tuple(reversed(sizes[b]))
        n = int(np.prod(shp))
        out[b] = [data[p + v * n: p + (v + 1) * n].reshape(shp) for v in range(nvars)]
        p += nvars * n
    assert p == data.size
    return out


@pytest.mark.parametrize("name,dim", [("case_cavity2d_16.json", 2),
                                      ("case_cavity3d_10_toporder.json", 3),
                                      ("case_lchannel_9.json", 3)])
def test_case_driver_json_matches_oracle(name, dim, tmp_path):
    path = os.path.join(GOLDEN, name)
    cfg = json.load(open(path))
    dump = str(tmp_path / "macros.bin")
    run(hipbuild.app_path("case_driver%dd" % dim), "Config=" + path, "Dump=" + dump)
    o = OracleCase(dim)
    steps, period, _ = cases.from_config(o, cfg)
    o.step(steps)
    o.update_macros()
    got = split_dump(dump, o.sizes, dim + 1)
    for b in o.sizes:
        for v in range(dim + 1):
            d = float(np.max(np.abs(got[b][v] - o.get_macro(b, v))))
            assert d == 0.0, (name, b, v, d)


def test_cavity3d_driver_matches_oracle(tmp_path):
    dump = str(tmp_path / "macros.bin")
    run(hipbuild.app_path("cavity3d"), "N=12", "Re=100", "Lid=0.05", "Steps=20",
        "Check=10", "Noise=1", "Dump=" + dump)
    o = OracleCase(3)
    cases.cavity3d(o, 12, re=100.0, ulid=0.05, noise=True)
    o.step(20)
    o.update_macros()
    got = split_dump(dump, o.sizes, 4)
    for v in range(4):
        assert np.array_equal(got[0][v], o.get_macro(0, v)), v


def test_cavity3d_driver_converges_with_residual(tmp_path):
    """Iterate(criteria, checkPeriod) through the C++ layer stops once the
    printed residual is below the criterion."""
    out = run(hipbuild.app_path("cavity3d"), "N=16", "Re=20", "Lid=0.05",
              "Criteria=1e-3", "Check=100")
    res = [float(l.split("=")[1]) for l in out.splitlines() if l.startswith("Residual of")]
    assert res, out[-500:]
    assert max(res[-4:]) < 1e-3
