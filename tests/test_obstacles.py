"""In-domain obstacles with half-way bounce-back (an extension: the
reference only skips ImmersedSolid nodes, model_kernel.inc:420,
scheme_kernel.inc:1418, and its bounce-back kernel is commented out,
boundary_kernel.inc:490-532).

The oracle's rule is pinned by
- Poiseuille flow between two obstacle plates: with BGK, half-way
  bounce-back puts the wall exactly half-way between nodes when
  tau = 1/2 + sqrt(3/16), so the profile equals the parabola
  u(y) = a/(2 nu) (y - y0)(y1 - y) with y0, y1 half a node outside the
  last fluid rows (a = the lattice acceleration, nu = (tau - 1/2)/3);
- exact mass conservation around obstacles in a periodic box, globally and
  inside a closed obstacle shell (no population crosses a wall).
"""
import math

import numpy as np
import pytest

from mplb_amd.enums import BoundaryScheme as BSch
from mplb_amd.enums import BoundarySurface as BS
from mplb_amd.enums import VertexType as VT
from oracle.oracle import OracleCase

CS = math.sqrt(3.0)
MAGIC_TAU = 0.5 + math.sqrt(3.0 / 16.0)


def plate_channel(be, nx=4, ny=24, tau_l=MAGIC_TAU, g=1e-6):
    """Periodic in x, EQMDiffuseRefl walls at Top/Bottom, obstacle plates on
    rows y = 2 and y = ny-3 across the whole width (through the periodic
    link), BodyForce_1st along x in code units (dt = 1)."""
    be.add_block(0, (nx, ny))
    be.add_connection(0, BS.Left, 0, BS.Right, VT.MDPeriodic)
    be.add_connection(0, BS.Right, 0, BS.Left, VT.MDPeriodic)
    be.set_collision(tau_l - 0.5, force_type=1, force_macros=True)
    be.add_boundary(0, BS.Left, BSch.None_, (), VT.MDPeriodic)
    be.add_boundary(0, BS.Right, BSch.None_, (), VT.MDPeriodic)
    be.add_boundary(0, BS.Top, BSch.EQMDiffuseRefl, (0.0, 0.0), VT.Wall)
    be.add_boundary(0, BS.Bottom, BSch.EQMDiffuseRefl, (0.0, 0.0), VT.Wall)
    solid = np.zeros((ny, nx), np.uint8)
    solid[2, :] = 1
    solid[ny - 3, :] = 1
    be.set_obstacles(0, solid)
    be.prepare()
    be.set_accel(0, (g, 0.0))
    z = np.zeros((ny, nx))
    be.set_macros(0, np.ones((ny, nx)), z, z)
    be.init_feq()
    be.set_dt(1.0)
    return solid


def poiseuille_error(u_code, ny, tau_l, g):
    u = u_code / CS                      # lattice velocity
    a = CS * g / 3.0                     # F_i = w rho (CS c).g = 3 w rho c.a
    nu = (tau_l - 0.5) / 3.0
    y0, y1 = 2.5, ny - 3.5
    ys = np.arange(3, ny - 3)
    ua = a / (2 * nu) * (ys - y0) * (y1 - ys)
    return float(np.max(np.abs(u[3:ny - 3] - ua)) / ua.max())


def test_plate_poiseuille_exact_at_magic_tau():
    o = OracleCase(2)
    plate_channel(o)
    o.step(80000)
    o.update_macros()
    u = o.get_macro(0, 1)
    # uniform along x between the plates (the outer strips see the corner
    # treatment of the EQM walls at the periodic faces)
    assert np.ptp(u[3:21, 0] - u[3:21, 2]) < 1e-18
    assert poiseuille_error(u[:, 1], 24, MAGIC_TAU, 1e-6) < 1e-9


def periodic_box_with_obstacles(be, dim, size, solid, seed=5):
    be.add_block(0, size)
    pairs = [(BS.Left, BS.Right), (BS.Right, BS.Left), (BS.Top, BS.Bottom),
             (BS.Bottom, BS.Top)]
    if dim == 3:
        pairs += [(BS.Front, BS.Back), (BS.Back, BS.Front)]
    for a, b in pairs:
        be.add_connection(0, a, 0, b, VT.MDPeriodic)
    be.set_collision(0.1, force_type=0)
    for a, _ in pairs:
        be.add_boundary(0, a, BSch.None_, (), VT.MDPeriodic)
    be.set_obstacles(0, solid)
    be.prepare()
    rng = np.random.default_rng(seed)
    shp = tuple(reversed(size))
    u = [0.05 * rng.standard_normal(shp) * (1 - solid) for _ in range(dim)]
    be.set_macros(0, 1 + 0.01 * rng.standard_normal(shp), *u)
    be.init_feq()
    be.set_dt(1.0)


def shell_masks(dim, n):
    """A closed shell (wall thickness 1) around a cube of side 4 inside a
    periodic box of side n; also returns the enclosed-node mask."""
    shp = (n,) * dim
    solid = np.zeros(shp, np.uint8)
    inner = np.zeros(shp, bool)
    sl_out = tuple(slice(3, 9) for _ in range(dim))
    sl_in = tuple(slice(4, 8) for _ in range(dim))
    solid[sl_out] = 1
    solid[sl_in] = 0
    inner[sl_in] = True
    solid[(slice(11, 13),) * dim] = 1  # a separate small block
    return solid, inner


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 14)])
def test_obstacles_conserve_mass_inside_and_out(dim, n):
    solid, inner = shell_masks(dim, n)
    o = OracleCase(dim)
    periodic_box_with_obstacles(o, dim, (n,) * dim, solid)
    f0 = o.get_f(0)
    m_in0 = float(f0[:, inner].sum())
    m0 = o.mass()
    o.step(300)
    f1 = o.get_f(0)
    assert abs(o.mass() - m0) / m0 < 1e-13
    assert abs(float(f1[:, inner].sum()) - m_in0) / m_in0 < 1e-13
    assert np.all(f1[:, solid.astype(bool)] == 0.0)  # solid nodes hold no f
    o.update_macros()
    assert np.max(np.abs(o.get_macro(0, 1))) > 1e-4   # the flow moved


def test_obstacle_next_to_a_wall_is_refused():
    o = OracleCase(2)
    o.add_block(0, (10, 10))
    o.set_collision(0.1)
    for s in ("Left", "Right", "Top", "Bottom"):
        o.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, (0.0, 0.0), VT.Wall)
    solid = np.zeros((10, 10), np.uint8)
    solid[1, 5] = 1  # its neighbour (5, 0) is a wall node
    o.set_obstacles(0, solid)
    with pytest.raises(RuntimeError, match="obstacle"):
        o.prepare()


def test_cavity_obstacles_plan_only():
    """The library refuses the same misplaced obstacle as the oracle, and
    reports obstacle nodes as ImmersedSolid."""
    from mplb_amd import cases, hip
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=-1)
    g.add_block(0, (10, 10))
    g.set_collision(0.1)
    for s in ("Left", "Right", "Top", "Bottom"):
        g.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, (0.0, 0.0), VT.Wall)
    solid = np.zeros((10, 10), np.uint8)
    solid[1, 5] = 1
    g.set_obstacles(0, solid)
    with pytest.raises(hip.MplbError, match="obstacle"):
        g.prepare()
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=-1)
    try:
        cases.cavity2d_obstacles(g, 20, [(5, 6, 4, 3)])
    except hip.MplbError:
        pass  # set_macros needs a device
    nt = g.node_info(0, 0)
    assert np.all(nt[6:9, 5:9] == VT.ImmersedSolid)
    assert nt[5, 5] == VT.Fluid


# ---- the HIP path against the oracle (FP64 arithmetic is bit-exact) ------------
def _gpu(dim):
    from mplb_amd import hip
    return hip.LatticeCase(dim, storage=hip.STORE_F64, device=0)


def _same(g, o, blocks=(0,)):
    g.update_macros()
    o.update_macros()
    for b in blocks:
        assert np.array_equal(g.get_f(b), o.get_f(b)), b
        for v in range(1 + o.dim):
            assert np.array_equal(g.get_macro(b, v), o.get_macro(b, v)), (b, v)


@pytest.mark.gpu
@pytest.mark.usefixtures("step_form")
def test_gpu_plate_channel_bitexact():
    print("obstacle plates through a periodic link, 600 steps", flush=True)
    o, g = OracleCase(2), _gpu(2)
    plate_channel(o)
    plate_channel(g)
    o.step(600)
    g.step(600)
    _same(g, o)


@pytest.mark.gpu
@pytest.mark.usefixtures("step_form")
@pytest.mark.parametrize("dim,n", [(2, 16), (3, 14)])
def test_gpu_obstacle_shell_bitexact(dim, n):
    print("obstacle shell %dD n=%d, 120 steps" % (dim, n), flush=True)
    solid, _ = shell_masks(dim, n)
    o, g = OracleCase(dim), _gpu(dim)
    periodic_box_with_obstacles(o, dim, (n,) * dim, solid)
    periodic_box_with_obstacles(g, dim, (n,) * dim, solid)
    o.step(120)
    g.step(120)
    _same(g, o)
    assert g.mass() == pytest.approx(o.mass(), rel=1e-13)


@pytest.mark.gpu
@pytest.mark.usefixtures("step_form")
def test_gpu_cavity_with_rectangles_bitexact_and_residual():
    from mplb_amd import cases
    rects = [(5, 6, 9, 4), (18, 20, 8, 9)]
    print("cavity 32^2 with rectangles %s" % rects, flush=True)
    o, g = OracleCase(2), _gpu(2)
    cases.cavity2d_obstacles(o, 32, rects, re=50.0)
    p = cases.cavity2d_obstacles(g, 32, rects, re=50.0)
    o.step(299)
    g.step(299)
    _same(g, o)
    o.step(1)
    g.step(1)
    o.update_macros()
    g.update_macros()
    ro, _ = o.residual(300)
    rg, _ = g.residual(300)
    assert rg == pytest.approx(ro, rel=1e-9)
    assert p["solid"].sum() > 0


@pytest.mark.gpu
@pytest.mark.usefixtures("step_form")
@pytest.mark.parametrize("storage,tol_u", [("f32", 2e-6), ("f16", 3e-4)])
def test_gpu_cavity_with_rectangles_reduced_storage(storage, tol_u):
    """The warm start's cavities (rectangles inside a lid-driven cavity) with
    FP32 f - w and FP16 storage, every step form (the bulk kernel's folded
    wall lanes beside obstacle links): velocity within tol_u of the FP64
    oracle after 300 steps (lid speed 0.087), mass to 1e-5."""
    from mplb_amd import cases, hip
    st = {"f32": hip.STORE_F32, "f16": hip.STORE_F16}[storage]
    rects = [(5, 6, 9, 4), (18, 20, 8, 9)]
    print("cavity 32^2 with rectangles, storage %s" % storage, flush=True)
    o = OracleCase(2)
    g = hip.LatticeCase(2, storage=st, device=0)
    cases.cavity2d_obstacles(o, 32, rects, re=50.0)
    cases.cavity2d_obstacles(g, 32, rects, re=50.0)
    o.step(300)
    g.step(300)
    o.update_macros()
    g.update_macros()
    du = max(float(np.max(np.abs(g.get_macro(0, v) - o.get_macro(0, v)))) for v in (1, 2))
    assert du < tol_u, du
    assert abs(g.mass() - o.mass()) < 1e-5 * o.mass()
