"""The learned warm start of the C5 configuration on held-out 128^2 cavities,
with the network trained on the GPU by the project's own kernels and its
weights kept in mplb_amd/data/unet128_c5.npz (written by
bench.py --workload warmstart --ws-save): the learned start must reach the
velocity residual 1e-8 in fewer steps than the rest start on the same
cases, and its converged velocity is compared with an FP64 run from rest."""
import os
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip, warmstart

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "mplb_amd", "data", "unet128_c5.npz")
SEED = 0x4D504C42
HELD_OUT = 16


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_learned_start_beats_rest_on_held_out_cases():
    net, n_train = warmstart.load_weights(WEIGHTS)
    n = net.n
    # the held-out cases follow the training range of case_params
    res, rects = warmstart.case_params(n_train + HELD_OUT, SEED, n=n)
    tr, tq = res[n_train:], rects[n_train:]
    log("warm start: %d held-out %d^2 cases, network trained on %d" % (HELD_OUT, n, n_train))
    g_rest, s_rest, _, t_rest = warmstart.solve(n, tr, tq, 1e-8)
    pred = warmstart.predict(net, warmstart.inputs(n, tr, tq))
    masks = [cases.rect_mask(n, r) for r in tq]
    init = warmstart.prediction_to_macros(pred, masks)
    g_warm, s_warm, _, t_warm = warmstart.solve(n, tr, tq, 1e-8, init=init)
    assert (s_rest > 0).all() and (s_warm > 0).all()
    ratio = s_rest.sum() / s_warm.sum()
    log("  steps: rest %d, learned %d (total ratio %.2f, per-case median %.2f); "
        "wall %.2f s vs %.2f s" % (s_rest.sum(), s_warm.sum(), ratio,
                                  np.median(s_rest / s_warm), t_rest, t_warm))
    assert ratio > 1.5
    assert np.median(s_rest / s_warm) > 1.5
    # the converged velocities against an FP64 run from rest to the same
    # residual (the FP64 lattice path is bit-exact with the CPU oracle)
    g64, s64, _, _ = warmstart.solve(n, tr, tq, 1e-8, storage=hip.STORE_F64)
    for g in (g_rest, g_warm, g64):
        g.update_macros()
    l2_warm = warmstart.velocity_l2(g_warm, g64, HELD_OUT)
    l2_rest = warmstart.velocity_l2(g_rest, g64, HELD_OUT)
    log("  velocity rel. L2 vs FP64 from rest: learned median %.3e max %.3e; "
        "rest (FP32) median %.3e max %.3e" % (np.median(l2_warm), l2_warm.max(),
                                              np.median(l2_rest), l2_rest.max()))
    # both stop at the residual criterion with the slowest transient still
    # present, so they differ from the FP64 run by that transient's share
    assert l2_warm.max() < 2e-2 and l2_rest.max() < 2e-2
    for g in (g_rest, g_warm, g64):
        g.close()
