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


def test_batch_c4_fp32_against_fp64():
    """The C4 batch (256 cavities, N in {64..256}, Re in [100,1000]) in FP32
    (one launch per step over all instances) against the same batch in
    FP64, whose every instance is bit-exact with its own oracle run
    (test_batch_f64_each_instance_bitexact): after 2,000 steps every
    instance's velocity within 1e-5 and density within 5e-6 (max-abs; the
    lid speed is 0.087; measured 4.1e-6 and 1.3e-6, the FP32 rounding of
    f - w accumulated over the run), and the batch's FP64 mass within 1e-5
    relative."""
    sizes, res = cases.batch_spec(256)
    steps = 2000
    g = hip.LatticeCase(2, storage=hip.STORE_F32, device=0)
    cases.batched_cavities(g, sizes, res)
    tot, ker = g.step_timed(steps)
    nodes = sum(n * n for n in sizes)
    log("C4 FP32: %d nodes, %.4f ms/step kernel, %.0f MLUPS" % (nodes, ker / steps,
                                                                 nodes / (ker / steps * 1e-3) / 1e6))
    r = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    cases.batched_cavities(r, sizes, res)
    r.step(steps)
    g.update_macros()
    r.update_macros()
    worst_u = worst_rho = 0.0
    for b in range(len(sizes)):
        du = max(float(np.max(np.abs(g.get_macro(b, v) - r.get_macro(b, v)))) for v in (1, 2))
        drho = float(np.max(np.abs(g.get_macro(b, 0) - r.get_macro(b, 0))))
        worst_u, worst_rho = max(worst_u, du), max(worst_rho, drho)
    dm = abs(g.mass() - r.mass()) / r.mass()
    log("C4 FP32 vs FP64 after %d steps: u %.3e, rho %.3e, mass %.3e rel"
        % (steps, worst_u, worst_rho, dm))
    assert worst_u < 1e-5 and worst_rho < 5e-6
    assert dm < 1e-5


def test_batch_restart_at_rest_replays_a_fresh_run():
    """The fallback restart: with a factor that any residual exceeds, every
    instance restarts at rest at its second check (step 2 * period) and then
    follows a fresh rest run exactly (FP64): it converges 2 * period steps
    later than the fresh run, with the same state and residual."""
    sizes, res = small_batch()
    crit, period = 1e-4, 50
    fresh = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    cases.batched_cavities(fresh, sizes, res)
    s_fresh, r_fresh, _ = fresh.iterate_batch(crit, period, 20000)
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    cases.batched_cavities(g, sizes, res)
    g.set_batch_restart(1e-3)
    steps, resid, _ = g.iterate_batch(crit, period, 20000)
    log("restarted steps %s fresh %s" % (list(steps), list(s_fresh)))
    assert np.all(s_fresh > 2 * period)
    assert list(g.batch_restarts()) == [1] * len(sizes)
    assert np.array_equal(steps, s_fresh + 2 * period)
    assert np.array_equal(resid, r_fresh)
    for b in range(len(sizes)):
        assert np.array_equal(g.get_f(b), fresh.get_f(b)), b
    # off by default and when set to 0
    g.set_batch_restart(0)
    g.iterate_batch(crit, period, 10 * period)
    assert list(g.batch_restarts()) == [0] * len(sizes)
    fresh.close()
    g.close()


def test_batch_compaction_matches_full_launches(monkeypatch):
    """Shrinking the step launches to the unconverged instances (and, below
    2^18 nodes, to one launch per step) changes neither any instance's
    convergence step nor its state, and the launches over every instance
    come back afterwards: later steps match too (FP64, bit-exact f)."""
    monkeypatch.delenv("MPLBHIP_FUSED", raising=False)
    sizes, res = cases.batch_spec(32, seed=777, sizes=(64, 96, 128, 160), re_range=(20.0, 300.0))
    log("compaction batch: %d nodes" % sum(n * n for n in sizes))
    assert sum(n * n for n in sizes) >= 1 << 18
    out = {}
    for compact in (0, 1):
        monkeypatch.setenv("MPLBHIP_BATCH_COMPACT", str(compact))
        g = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
        cases.batched_cavities(g, sizes, res)
        steps, resid, done = g.iterate_batch(1e-4, 50, 50000)
        g.step(7)
        out[compact] = (steps, resid, done, [g.get_f(b) for b in range(len(sizes))])
        g.close()
    (s0, r0, d0, f0), (s1, r1, d1, f1) = out[0], out[1]
    log("steps %s" % list(s1))
    assert np.all(s0 > 0)
    assert np.array_equal(s0, s1) and d0 == d1
    # the per-block partial sums are grouped differently in the two forms
    assert np.allclose(r0, r1, rtol=1e-12, atol=0)
    for b in range(len(sizes)):
        assert np.array_equal(f0[b], f1[b]), b


MIXED = [("cavity", 16, 50.0), ("channel", [12], 9, 0.9, 2e-5),
         ("cavity", 20, 80.0), ("channel", [7, 6], 11, 0.8, 1e-5)]


def _oracle_instance(inst):
    o = OracleCase(2)
    if inst[0] == "cavity":
        p = cases.cavity2d(o, inst[1], re=inst[2], ulid=0.05)
    else:
        p = cases.channel_multi(o, inst[1], inst[2], inst[3], inst[4])
    return o, p["dt"]


def test_batch_mixed_cavities_and_channels():
    """Cavities and force-driven periodic channels (one self-linked block,
    and a channel of two linked blocks) in one batch: each instance
    converges at the check step where its own velocity residual, summed over
    its blocks, first meets the criterion, exactly as a single run of it on
    the oracle, and keeps the state it converged with (bit for bit)."""
    crit, period, limit = 1e-4, 50, 40000
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    insts = cases.batched_mixed(g, MIXED)
    steps, resid, done = g.iterate_batch(crit, period, limit)
    log("mixed batch steps %s done %d" % (list(steps), done))
    assert np.all(steps > 0)
    for inst, spec in zip(insts, MIXED):
        o, dt = _oracle_instance(spec)
        t, r = 0, None
        while t < limit:
            o.step(period)
            t += period
            o.update_macros()
            _, ratios = o.residual(period)
            r = max(ratios[1], ratios[2]) / (period * dt)
            if r < crit:
                break
        for k, b in enumerate(inst["blocks"]):
            assert steps[b] == t, (spec, b, steps[b], t)
            assert abs(resid[b] - r) <= 1e-9 * r
            assert np.array_equal(g.get_f(b), o.get_f(k)), (spec, b)
    # the channels did move and the instances took different times
    assert len(set(int(s) for s in steps)) > 1
