"""Batched independent cavities (many blocks, one launch per step), each
block with its own relaxation time and time step, against single-case
oracle runs; per-instance convergence with freezing."""
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip
from oracle.oracle import OracleCase

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("step_form")]


def log(m):
    print(m, flush=True)


def small_batch():
    return cases.batch_spec(7, seed=12345, sizes=(16, 20, 24, 32), re_range=(20.0, 200.0))


def test_batch_f64_each_instance_bitexact():
    sizes, res = small_batch()
    log("batch F64 sizes=%s" % sizes)
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    cases.batched_cavities(g, sizes, res)
    g.step(150)
    g.update_macros()
    for b, (n, re) in enumerate(zip(sizes, res)):
        o = OracleCase(2)
        cases.cavity2d(o, n, re=re, ulid=0.05)
        o.step(150)
        o.update_macros()
        assert np.array_equal(g.get_f(b), o.get_f(0)), b
        for v in range(3):
            assert np.array_equal(g.get_macro(b, v), o.get_macro(0, v)), (b, v)


def _oracle_velocity_convergence(n, re, crit, period, limit):
    o = OracleCase(2)
    p = cases.cavity2d(o, n, re=re, ulid=0.05)
    t = 0
    while t < limit:
        o.step(period)
        t += period
        o.update_macros()
        _, ratios = o.residual(period)
        r = max(ratios[1], ratios[2]) / (period * p["dt"])
        if r < crit:
            return t, r
    return -1, r


def test_batch_converges_per_instance_like_single_runs():
    """Each instance stops at the check step where its own FP64 velocity
    residual first drops below the criterion, as a single run would; the
    frozen instances keep their converged state."""
    sizes, res = small_batch()
    crit, period = 1e-4, 50
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    cases.batched_cavities(g, sizes, res)
    steps, resid, done = g.iterate_batch(crit, period, 20000)
    log("batch steps %s done %d" % (list(steps), done))
    assert np.all(steps > 0)
    assert done == max(steps)
    for b, (n, re) in enumerate(zip(sizes, res)):
        t, r = _oracle_velocity_convergence(n, re, crit, period, 20000)
        assert steps[b] == t, (b, steps[b], t)
        assert abs(resid[b] - r) <= 1e-9 * r
    # a frozen instance holds the state it converged with
    o = OracleCase(2)
    cases.cavity2d(o, sizes[0], re=res[0], ulid=0.05)
    o.step(int(steps[0]))
    assert np.array_equal(g.get_f(0), o.get_f(0))


def test_batch_c4_fp32_runs():
    """The C4 batch (256 cavities, N in {64..256}, Re in [100,1000]) in FP32:
    one launch per step over all instances; a short run stays finite."""
    sizes, res = cases.batch_spec(256)
    g = hip.LatticeCase(2, storage=hip.STORE_F32, device=0)
    cases.batched_cavities(g, sizes, res)
    tot, ker = g.step_timed(100)
    nodes = sum(n * n for n in sizes)
    log("C4 FP32: %d nodes, %.4f ms/step kernel, %.0f MLUPS" % (nodes, ker / 100,
                                                                 nodes / (ker / 100 * 1e-3) / 1e6))
    g.update_macros()
    for b in (0, 100, 255):
        assert np.all(np.isfinite(g.get_macro(b, 1)))
