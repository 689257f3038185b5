"""HIP path against the CPU oracle, through the C ABI.

FP64 storage reproduces the reference arithmetic operation for operation,
so populations and macros must agree bit for bit (the acceptance bar is
max-abs 1e-12 on f and macros, 1e-13 relative on the total mass).  Every
test here runs FP64 storage; the FP32 and FP16 forms are checked against the
FP64 oracle with stated tolerances in test_precision_gpu.py (whole-lattice
launch) and test_reduced_split_gpu.py (every split-step form, linked blocks).
"""
import math
import os
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip
from mplb_amd.enums import BoundaryScheme as BSch
from mplb_amd.enums import BoundarySurface as BS
from mplb_amd.enums import VertexType as VT
from oracle.oracle import OracleCase

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("step_form")]

F_TOL = 1e-12
MASS_TOL = 1e-13


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair(dim, build, storage=hip.STORE_F64):
    g = hip.LatticeCase(dim, storage=storage, device=0)
    pg = build(g)
    o = OracleCase(dim)
    po = build(o)
    return g, o, pg


def compare(g, o, what, f_tol=F_TOL, macro_tol=F_TOL, mass_tol=MASS_TOL):
    worst = 0.0
    for b in g.sizes:
        fg, fo = g.get_f(b), o.get_f(b)
        d = float(np.max(np.abs(fg - fo)))
        worst = max(worst, d)
        assert d <= f_tol, "%s block %d: max|f_gpu - f_oracle| = %g" % (what, b, d)
    g.update_macros()
    o.update_macros()
    for b in g.sizes:
        for v in range(1 + g.dim):
            d = float(np.max(np.abs(g.get_macro(b, v) - o.get_macro(b, v))))
            assert d <= macro_tol, "%s block %d var %d: %g" % (what, b, v, d)
    mg, mo = g.mass(), o.mass()
    assert abs(mg - mo) / abs(mo) <= mass_tol, (what, mg, mo)
    return worst


@pytest.mark.parametrize("steps", [1, 7, 60])
def test_f64_cavity2d_bitexact(steps):
    log("cavity2d 16^2 F64 steps=%d" % steps)
    g, o, _ = pair(2, lambda be: cases.cavity2d(be, 16, re=100.0))
    g.step(steps)
    o.step(steps)
    assert compare(g, o, "cavity2d") == 0.0


def test_f64_initial_state_and_restart():
    log("cavity2d 10^2 F64 t=0 and set_f restart")
    g, o, _ = pair(2, lambda be: cases.cavity2d(be, 10, re=50.0))
    assert compare(g, o, "t=0") == 0.0
    o.step(5)
    f5 = o.get_f(0)
    g.set_f(0, f5)
    o2 = OracleCase(2)
    cases.cavity2d(o2, 10, re=50.0)
    o2.set_f(0, f5)
    g.step(9)
    o2.step(9)
    assert compare(g, o2, "restart") == 0.0


def test_f64_ragged_2d_and_order():
    """non-square block, BC order Top first (the JSON order)"""
    log("ragged 24x20 F64")

    def build(be):
        be.add_block(0, (24, 20))
        mesh, dt, tau_l = cases.cavity_params(24, 400.0, 0.08)
        be.set_collision(cases.tau_ref_for(tau_l, dt))
        for s, v in (("Top", (0.08 * cases.CS, 0.0)), ("Bottom", (0.0, 0.0)),
                     ("Left", (0.0, 0.01)), ("Right", (0.0, 0.0))):
            be.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, v, VT.Wall)
        be.prepare()
        shp = (20, 24)
        be.set_macros(0, cases.density_noise(shp, amplitude=1e-3),
                      np.zeros(shp), np.zeros(shp))
        be.init_feq()
        be.set_dt(dt)
    g, o, _ = pair(2, build)
    g.step(41)
    o.step(41)
    assert compare(g, o, "ragged") == 0.0


@pytest.mark.parametrize("steps", [1, 25])
def test_f64_cavity3d_bitexact(steps):
    log("cavity3d 12^3 F64 noise steps=%d" % steps)
    g, o, _ = pair(3, lambda be: cases.cavity3d(be, 12, re=100.0, noise=True))
    g.step(steps)
    o.step(steps)
    assert compare(g, o, "cavity3d") == 0.0


def test_f64_box3d_nonuniform():
    log("box 13x9x11 F64, JSON BC order")
    g, o, _ = pair(3, lambda be: cases.cavity3d(
        be, 11, re=80.0, noise=True, size=(13, 9, 11),
        order=("Top", "Bottom", "Left", "Right", "Front", "Back")))
    g.step(30)
    o.step(30)
    assert compare(g, o, "box3d") == 0.0


def test_f64_row_workgroups_bitexact(monkeypatch):
    """The bulk kernel's row-workgroup shape (four 64-node x tiles of one row
    per workgroup), which needs rows of a multiple of 256 nodes: a 256 x 8 x
    12 box, split step, against the oracle."""
    monkeypatch.setenv("MPLBHIP_FUSED", "0")
    monkeypatch.setenv("MPLBHIP_XGROUP", "2")
    log("box 256x8x12 F64, row workgroups")
    g, o, _ = pair(3, lambda be: cases.cavity3d(be, 12, re=150.0, noise=True,
                                                size=(256, 8, 12)))
    g.step(9)
    o.step(9)
    assert compare(g, o, "rowwg") == 0.0


def test_f64_lchannel_multiblock_bitexact():
    """three blocks, VirtualBoundary halos, extrapolation outlet"""
    log("lchannel n=9 F64")
    g, o, _ = pair(3, lambda be: cases.lchannel(be, n=9, top_lid=-0.05))
    g.step(40)
    o.step(40)
    assert compare(g, o, "lchannel") == 0.0


def test_f64_periodic_force_bitexact():
    log("periodic 3D with body force F64")
    rng = np.random.default_rng(5)
    size = (10, 8, 6)
    shp = tuple(reversed(size))
    rho = 1 + 0.01 * rng.standard_normal(shp)
    u = [0.02 * rng.standard_normal(shp) for _ in range(3)]
    g, o, _ = pair(3, lambda be: cases.periodic_box(be, size, 0.7, rho, *u,
                                                     force=(1e-4, -2e-4, 5e-5)))
    g.step(33)
    o.step(33)
    assert compare(g, o, "periodic") == 0.0


def test_f64_channel2d_force_bitexact():
    log("channel2d with force F64")
    g, o, _ = pair(2, lambda be: cases.channel2d(be, 7, 13, 0.9, 1e-5))
    g.step(120)
    o.step(120)
    assert compare(g, o, "channel") == 0.0


def test_f64_residual_matches_oracle():
    log("residual F64")
    g, o, _ = pair(2, lambda be: cases.cavity2d(be, 20, re=100.0))
    for _ in range(3):
        g.step(50)
        o.step(50)
        rg, pg = g.residual(50)
        o.update_macros()
        ro, po = o.residual(50)
        assert math.isclose(rg, ro, rel_tol=1e-10), (rg, ro)
        assert np.allclose(pg, po, rtol=1e-10, atol=0)


def test_c1_cavity2d_128_full_run():
    """Config 1 in full: D2Q9 FP64, 128^2, Re=100, lid 0.05, 10,000 steps."""
    log("C1 cavity2d 128^2 F64 10000 steps")
    g, o, _ = pair(2, lambda be: cases.cavity2d(be, 128, re=100.0, ulid=0.05))
    g.step(10000)
    o.step(10000)
    d = compare(g, o, "C1")
    log("C1 max|df| = %g" % d)
    assert d == 0.0


def test_c2_cavity3d_full_size_few_steps():
    """Config 2 at its full size, 256^3 FP64 with density noise, for as many
    steps as the oracle finishes quickly; the long run is covered by
    size-independent checks in test_precision_gpu.py."""
    n = int(os.environ.get("MPLB_C2_N", "256"))
    steps = int(os.environ.get("MPLB_C2_STEPS", "3"))
    log("C2 cavity3d %d^3 F64 %d steps" % (n, steps))
    g, o, _ = pair(3, lambda be: cases.cavity3d(be, n, re=400.0, ulid=0.05, noise=True))
    g.step(steps)
    o.step(steps)
    assert compare(g, o, "C2") == 0.0


def test_f64_iterate_to_convergence_matches_oracle():
    """Iterate(criteria, checkPeriod) (evolution.cpp:102-135): the fused
    check step produces the same residual sequence and stopping step as the
    oracle's UpdateMacroVars + CalcResidualError loop."""
    log("iterate to residual 1e-4, 2D 24^2")
    period, crit = 200, 1e-4
    g, o, _ = pair(2, lambda be: cases.cavity2d(be, 24, re=20.0, ulid=0.05))
    steps, rg = g.iterate(crit, period, 200000)
    it, ro = 0, 1.0
    while ro >= crit:
        o.step(period)
        it += period
        o.update_macros()
        ro, _ = o.residual(period)
    assert steps == it, (steps, it)
    assert math.isclose(rg, ro, rel_tol=1e-9), (rg, ro)
    assert compare(g, o, "iterate") == 0.0


def test_f64_iterate_transient_refreshes_macros():
    """Iterate(steps, checkPeriod): macros are those of the last check step."""
    log("iterate transient")
    g, o, _ = pair(3, lambda be: cases.cavity3d(be, 9, re=50.0, noise=True))
    steps, res = g.iterate(-1.0, 10, 35)
    assert steps == 35 and res == -1.0
    o.step(30)
    o.update_macros()
    for v in range(4):
        assert np.array_equal(g.get_macro(0, v), o.get_macro(0, v))
    o.step(5)
    assert np.array_equal(g.get_f(0), o.get_f(0))


@pytest.mark.parametrize("dim,axis,sizes", [
    (2, 0, [(4, 5), (3, 5)]),
    (3, 0, [(3, 4, 5), (4, 4, 5)]),
    (3, 1, [(5, 3, 4), (5, 4, 4)]),
    (3, 2, [(4, 5, 3), (4, 5, 3)]),
])
def test_f64_split_box_junction_halos_bitexact(dim, axis, sizes):
    """A periodic box cut into two blocks joined by VirtualBoundary
    connections (the junction rims whose halo cells OPS's group order
    leaves one exchange old, pinned in test_oracle_pins.py), 30 steps."""
    from tests.test_oracle_pins import _build_split_box
    from mplb_amd import lattice
    log("split box dim=%d axis=%d %s" % (dim, axis, sizes))
    rng = np.random.default_rng(5 + axis)
    glob = list(sizes[0])
    glob[axis] += sizes[1][axis]
    Q = 9 if dim == 2 else 19
    _, w = lattice.lattice(dim)
    f_glob = w.reshape((Q,) + (1,) * dim) * (1 + 0.2 * rng.random((Q,) + tuple(reversed(glob))))
    g = hip.LatticeCase(dim, storage=hip.STORE_F64, device=0)
    _build_split_box(g, dim, sizes, axis, 0.8, f_glob)
    o = OracleCase(dim)
    _build_split_box(o, dim, sizes, axis, 0.8, f_glob)
    g.step(30)
    o.step(30)
    assert compare(g, o, "split box") == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_f64_residual_closed_form(dim):
    """The uniform forced periodic flow whose residual ratios have a closed
    form (test_oracle_pins.test_residual_closed_form), on the device."""
    from tests.test_oracle_pins import _uniform_forced_box
    log("residual closed form dim=%d" % dim)
    size = (6, 5) if dim == 2 else (6, 5, 4)
    u0 = (0.02, -0.013, 0.011)[:dim]
    gv = (1e-4, 2e-4, -3e-4)[:dim]
    dt, tau_l, P = 0.37, 0.8, 7
    g = hip.LatticeCase(dim, storage=hip.STORE_F64, device=0)
    _uniform_forced_box(g, dim, size, tau_l, dt, u0, gv)
    g.update_macros()
    g.residual(P)
    g.step(P)
    res, per = g.residual(P)
    exp = [(P * dt * gv[d]) ** 2 / (u0[d] + (P + 0.5) * dt * gv[d]) ** 2 for d in range(dim)]
    assert per[0] < 1e-28
    assert np.allclose(per[1:], exp, rtol=1e-9, atol=0), (per, exp)
    assert math.isclose(res, max(exp) / (P * dt), rel_tol=1e-9)
