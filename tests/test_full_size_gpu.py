"""Cases at or near their benchmark size against the FP64 oracle, in the
launch shape the autotuner picks: the first 40 steps of the 256 batched 2D
cavities (7.3M nodes in one launch per step), each instance bit for bit
against a single oracle run of its cavity; the swap scheme on a 128^3
cavity."""
import numpy as np
import pytest

from mplb_amd import cases, hip
from oracle.oracle import OracleCase

pytestmark = pytest.mark.gpu


def test_c4_batch256_fp64_first_steps_bitexact():
    steps = 40  # one replay of the 32-step graph, then single steps
    sizes, res = cases.batch_spec(256)
    print("C4: 256 cavities, %d nodes, FP64, %d steps" % (sum(n * n for n in sizes), steps),
          flush=True)
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    cases.batched_cavities(g, sizes, res)
    g.step(steps)
    fold, xgroup, h2, bulk_nodes = g.launch_shape()
    print("launch shape: fold %d, xgroup %d, bulk nodes %d" % (fold, xgroup, bulk_nodes),
          flush=True)
    g.update_macros()
    for b, (n, re) in enumerate(zip(sizes, res)):
        o = OracleCase(2)
        cases.cavity2d(o, n, re=re, ulid=0.05)
        o.step(steps)
        o.update_macros()
        assert np.array_equal(g.get_f(b), o.get_f(0)), b
        for v in range(3):
            assert np.array_equal(g.get_macro(b, v), o.get_macro(0, v)), (b, v)
    g.close()


def test_swap_scheme_cavity_128_fp64_bitexact():
    """The swap scheme (Apps/3DCavity-Swap) on a 128^3 cavity, 2.1M nodes,
    for 12 steps: every population and macro bit for bit."""
    n, steps = 128, 12
    print("swap scheme cavity %d^3 FP64, %d steps" % (n, steps), flush=True)
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
    cases.cavity3d_swap(g, n)
    o = OracleCase(3)
    cases.cavity3d_swap(o, n)
    g.step(steps)
    o.step(steps)
    assert np.array_equal(g.get_f(0), o.get_f(0))
    g.update_macros()
    o.update_macros()
    for v in range(4):
        assert np.array_equal(g.get_macro(0, v), o.get_macro(0, v)), v
    assert abs(g.mass() - o.mass()) <= 1e-13 * abs(o.mass())
    g.close()
