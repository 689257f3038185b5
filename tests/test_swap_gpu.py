"""The swap scheme (Scheme_StreamCollision_Swap, Src/evolution.cpp:195-252)
and FDPeriodic boundaries on the GPU against the CPU oracle, FP64 storage:
populations, macros and mass must agree bit for bit (mass to 1e-13, its
reduction order differs)."""
import math
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip
from oracle.oracle import OracleCase

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(g, o, what):
    for b in g.sizes:
        assert np.array_equal(g.get_f(b), o.get_f(b)), (what, b)
    g.update_macros()
    o.update_macros()
    for b in g.sizes:
        for v in range(4):
            assert np.array_equal(g.get_macro(b, v), o.get_macro(b, v)), (what, b, v)
    assert abs(g.mass() - o.mass()) <= 1e-13 * abs(o.mass()), what


@pytest.mark.parametrize("steps", [1, 30])
def test_swap_cavity_bitexact(steps):
    log("swap cavity 12^3 F64 %d steps" % steps)
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
    cases.cavity3d_swap(g, 12, re=100.0, noise=True)
    o = OracleCase(3)
    cases.cavity3d_swap(o, 12, re=100.0, noise=True)
    g.step(steps)
    o.step(steps)
    compare(g, o, "swap cavity")


def test_swap_box_nonuniform_json_order():
    log("swap cavity 13x9x11, Top first")
    build = lambda be: cases.cavity3d_swap(be, 11, re=80.0, noise=True, size=(13, 9, 11),  # noqa
                                           order=("Top", "Bottom", "Left", "Right", "Front",
                                                  "Back"))
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
    build(g)
    o = OracleCase(3)
    build(o)
    g.step(40)
    o.step(40)
    compare(g, o, "swap box")


@pytest.mark.parametrize("fd_axis,force", [(None, None), (None, (1e-4, -2e-4, 5e-5)),
                                           (0, None), (1, (2e-5, 0.0, 0.0)), (2, None)])
def test_swap_periodic_and_fdperiodic_bitexact(fd_axis, force):
    """Self-linked faces (MDPeriodic halos of f, BodyForce_1st_Swap) and
    FDPeriodic faces (KerCutCellPeriodic3D copying the exchanged halo)."""
    log("swap box fd_axis=%s force=%s" % (fd_axis, force))
    rng = np.random.default_rng(17)
    size = (9, 8, 7)
    shp = tuple(reversed(size))
    rho = 1 + 0.01 * rng.standard_normal(shp)
    u = [0.02 * rng.standard_normal(shp) for _ in range(3)]
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
    cases.swap_box(g, size, 0.8, rho, *u, force=force, fd_axis=fd_axis)
    o = OracleCase(3)
    cases.swap_box(o, size, 0.8, rho, *u, force=force, fd_axis=fd_axis)
    g.step(25)
    o.step(25)
    compare(g, o, "swap box")


def test_swap_iterate_residual_matches_oracle():
    """Iterate(criteria, period) on the swap scheme: the check step measures
    the macros of its result; same residual sequence and stopping step."""
    log("swap iterate to 1e-3")
    period, crit = 20, 1e-3
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
    cases.cavity3d_swap(g, 10, re=50.0, noise=False)
    steps, rg = g.iterate(crit, period, 20000)
    o = OracleCase(3)
    cases.cavity3d_swap(o, 10, re=50.0, noise=False)
    it, ro = 0, 1.0
    while ro >= crit:
        o.step(period)
        it += period
        o.update_macros()
        ro, _ = o.residual(period)
    assert steps == it, (steps, it)
    assert math.isclose(rg, ro, rel_tol=1e-9), (rg, ro)
    compare(g, o, "swap iterate")
