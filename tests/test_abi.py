"""CPU-side checks of the C ABI: the library loads, exports every entry
point the header declares, and its plan (node typing, geometry, boundary
programs) reproduces the oracle's per-node tagging.  No compute calls."""
import ctypes

import numpy as np
import pytest

from mplb_amd import cases, hip
from mplb_amd.enums import BoundaryScheme as BSch
from mplb_amd.enums import BoundarySurface as BS
from mplb_amd.enums import VertexType as VT
from oracle.oracle import OracleCase


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hip.LIB_PATH)
    names = hip.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in hip.load().mplbhip_version()


class PlanOnly(hip.LatticeCase):
    """device -1: everything up to Partition; state calls are skipped."""

    def __init__(self, dim):
        super().__init__(dim, device=-1)

    def set_macros(self, *a, **k):
        pass

    def init_feq(self):
        pass

    def set_dt(self, dt):
        pass

    def set_accel(self, block, g):
        pass


def _compare_tagging(build):
    dims = {}
    p = None
    for Backend in ("plan", "oracle"):
        if Backend == "plan":
            dim, fn = build
            p = PlanOnly(dim)
            fn(p)
        else:
            dim, fn = build
            o = OracleCase(dim)
            fn(o)
    for b in p.sizes:
        for which in (0, 1):
            assert np.array_equal(p.node_info(b, which), o.get_int(b, which)), (b, which)
    return p


CASES = {
    "cavity2d": (2, lambda be: cases.cavity2d(be, 12)),
    "cavity2d_ragged": (2, lambda be: cases.cavity2d(be, 9, order=("Top", "Bottom", "Left", "Right"))),
    "cavity3d": (3, lambda be: cases.cavity3d(be, 7, noise=False)),
    "cavity3d_json_order": (3, lambda be: cases.cavity3d(
        be, 6, noise=False, order=("Top", "Bottom", "Left", "Right", "Front", "Back"))),
    "box3d": (3, lambda be: cases.cavity3d(be, 5, noise=False, size=(9, 6, 5))),
    "lchannel": (3, lambda be: cases.lchannel(be, n=7)),
    "channel2d": (2, lambda be: cases.channel2d(be, 6, 9, 0.8, 1e-6)),
    "periodic3d": (3, lambda be: cases.periodic_box(
        be, (5, 4, 6), 0.8, np.ones((6, 4, 5)), np.zeros((6, 4, 5)),
        np.zeros((6, 4, 5)), np.zeros((6, 4, 5)))),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_plan_tagging_matches_oracle(name):
    _compare_tagging(CASES[name])


def test_edge_surfaces_and_topback_quirk():
    """Edge-surface boundaries (RightTop, TopBack ...) and the TopBack y
    range quirk of block.cpp:35 tag the same nodes in plan and oracle."""
    def build(be):
        be.add_block(0, (6, 5, 7))
        be.set_collision(0.1)
        wall = (0.0, 0.0, 0.0)
        for s in ("Left", "Right", "Top", "Bottom", "Front", "Back"):
            be.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, wall, VT.Wall)
        be.add_boundary(0, BS.RightTop, BSch.EQMDiffuseRefl, (0.01, 0, 0), VT.Inlet)
        be.add_boundary(0, BS.TopBack, BSch.None_, (), VT.Symmetry)
        be.add_boundary(0, BS.LeftFront, BSch.None_, (), VT.OutLet)
        be.prepare()
    _compare_tagging((3, build))


def test_plan_only_context_refuses_compute():
    p = PlanOnly(2)
    cases.cavity2d(p, 8)
    with pytest.raises(hip.MplbError, match="without a device"):
        p.step(1)


def test_unsupported_scheme_is_an_error():
    """BounceBack has no active reference kernel; FDPeriodic copies f's halo,
    which only the swap scheme exchanges (refused at prepare under
    stream-collision); the swap scheme itself is D3Q19 FP64 only."""
    p = PlanOnly(2)
    p.add_block(0, (8, 8))
    p.set_collision(0.1)
    with pytest.raises(hip.MplbError, match="not implemented"):
        p.add_boundary(0, BS.Left, BSch.BounceBack, (), VT.Wall)
    p.add_boundary(0, BS.Left, BSch.FDPeriodic, (), VT.FDPeriodic)
    with pytest.raises(hip.MplbError, match="Swap"):
        p.prepare()
    q = PlanOnly(2)
    with pytest.raises(hip.MplbError, match="D3Q19"):
        q.set_scheme(11)
    r = hip.LatticeCase(3, storage=hip.STORE_F32, device=-1)
    with pytest.raises(hip.MplbError, match="FP64"):
        r.set_scheme(11)


def test_define_order_errors():
    p = PlanOnly(3)
    with pytest.raises(hip.MplbError, match="id order"):
        p.add_block(1, (8, 8, 8))
    p.add_block(0, (8, 8, 8))
    with pytest.raises(hip.MplbError, match="< 3 nodes"):
        p.add_block(1, (8, 2, 8))
    p.add_connection(0, BS.Left, 0, BS.Right, VT.MDPeriodic)
    with pytest.raises(hip.MplbError, match="already defined"):
        p.add_connection(0, BS.Left, 0, BS.Right, VT.MDPeriodic)


@pytest.mark.parametrize("fold", ["0", "1", "2", "3"])
def test_launch_shape_counts_bulk_nodes(monkeypatch, fold):
    """The nodes one bulk launch updates: the interior of planes 1 .. n-2 at
    fold level 0; plus the x-face walls (simple EQM regions) at level 1;
    at level 2 at most whole planes; at level 3 every node but the eight
    corners (whose wall scheme keeps state between steps)."""
    monkeypatch.setenv("MPLBHIP_XFACE_FOLD", fold)
    g = hip.LatticeCase(3, storage=hip.STORE_F32, device=-1)
    nx, ny, nz = 20, 16, 12
    g.add_block(0, (nx, ny, nz))
    g.set_collision(1.0, force_type=0)
    for s in ("Left", "Right", "Top", "Bottom", "Front", "Back"):
        g.add_boundary(0, cases.BS[s], BSch.EQMDiffuseRefl, (0.0, 0.0, 0.0), VT.Wall)
    g.prepare()
    f, _, h2, nodes = g.launch_shape()
    assert f == int(fold) and h2 == 0
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    if fold == "0":
        assert nodes == interior
    elif fold == "1":
        assert nodes == nx * (ny - 2) * (nz - 2)
    elif fold == "2":
        assert nx * (ny - 2) * (nz - 2) <= nodes <= nx * ny * (nz - 2)
    else:
        assert nodes == nx * ny * nz - 8


def test_lattice_and_collision_type_arguments():
    """define_lattice takes 9 in 2D and 19 or 15 in 3D; the thermal
    collision needs FP64 storage; the D3Q15 plan tags nodes as the oracle
    does."""
    with pytest.raises(hip.MplbError, match="d3q15"):
        hip.LatticeCase(2, device=-1, lattice=19)
    with pytest.raises(hip.MplbError, match="d3q15"):
        hip.LatticeCase(3, device=-1, lattice=27)
    r = hip.LatticeCase(3, storage=hip.STORE_F16, device=-1, lattice=15)
    with pytest.raises(hip.MplbError, match="FP64"):
        r.set_collision_type(1)
    with pytest.raises(hip.MplbError, match="collision type 2"):
        r.set_collision_type(2)

    class PlanOnly15(PlanOnly):
        def __init__(self, dim):
            hip.LatticeCase.__init__(self, dim, device=-1, lattice=15)

    p = PlanOnly15(3)
    cases.cavity3d(p, 7, noise=False)
    o = OracleCase(3, lattice=15)
    cases.cavity3d(o, 7, noise=False)
    for which in (0, 1):
        assert np.array_equal(p.node_info(0, which), o.get_int(0, which))


def test_folded_region_shift_faces():
    """The faces across which a node of a simple region streams nothing
    (node_info which=3: bit 2a low side, 2a+1 high side of axis a), whose
    sources the bulk kernel moves back inside the block: every wall face
    of a closed cavity at its face, edge and corner nodes; none inside or
    on periodic faces; and a region whose program keeps state (the cavity's
    corners persist populations) is not simple."""
    p = PlanOnly(3)
    cases.cavity3d(p, 7, noise=False)
    simple, adj = p.node_info(0, 2), p.node_info(0, 3)   # [k][j][i]
    n = 7
    for k in range(n):
        for j in range(n):
            for i in range(n):
                want = 0
                for a, c in enumerate((i, j, k)):
                    if c == 0:
                        want |= 1 << (2 * a)
                    elif c == n - 1:
                        want |= 1 << (2 * a + 1)
                corner = sum(c in (0, n - 1) for c in (i, j, k)) == 3
                if corner:
                    assert simple[k, j, i] == 0 and adj[k, j, i] == 0, (i, j, k)
                else:
                    assert simple[k, j, i] == 1 and adj[k, j, i] == want, (i, j, k)

    q = PlanOnly(3)
    cases.periodic_box(q, (5, 4, 6), 0.8, np.ones((6, 4, 5)), np.zeros((6, 4, 5)),
                       np.zeros((6, 4, 5)), np.zeros((6, 4, 5)))
    assert np.all(q.node_info(0, 3) == 0)
    assert np.all(q.node_info(0, 2) == 1)
    with pytest.raises(hip.MplbError, match="which"):
        q.node_info(0, 4)
