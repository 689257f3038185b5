"""C2 at its stated length in FP64: the 3D lid-driven cavity, D3Q19 256^3,
Re 400, lid 0.05, seeded density noise, 2,000 steps (62 replays of the
32-step graph after the autotuned first step).  The populations must equal
the FP64 oracle's bit for bit; the oracle's state is committed as a digest
(tests/golden/c2_fp64_digest.json, written by make_c2_digest.py in the same
directory): SHA-256 of f in [q][k][j][i] order and probe values, bit for
bit, and the mass to 1e-11 relative (the committed mass is a plain
running sum of the 3.2e8 populations, written before the oracle's mass
became a compensated sum; the device's tree of partial sums is 1.2e-13
from it)."""
import json
import os
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip

from .golden.make_c2_digest import PROBES, digest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "c2_fp64_digest.json")


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="C2 digest not generated")
def test_c2_fp64_2000_steps_digest():
    ref = json.load(open(GOLDEN))
    n, steps = ref["n"], ref["steps"]
    print("C2 FP64 %d^3, %d steps" % (n, steps), flush=True)
    sys.stdout.flush()
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
    cases.cavity3d(g, n, re=400.0, ulid=0.05, noise=True)
    g.step(steps)
    f = g.get_f(0)
    for p, v in ref["probes"]:
        assert f[tuple(p)] == float(v), (p, repr(f[tuple(p)]), v)
    assert digest(f) == ref["sha256_f"]
    m = g.mass()
    assert abs(m - float(ref["mass"])) <= 1e-11 * abs(float(ref["mass"])), (m, ref["mass"])
    assert np.isfinite(f).all()
